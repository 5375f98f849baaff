// kernel.cpp -- per-shot time loops (see kernel.hpp).
#include "kernel.hpp"

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>
#include <stdexcept>

namespace rtm3d {

namespace {

double now_s()
{
    using namespace std::chrono;
    return duration_cast<microseconds>(steady_clock::now().time_since_epoch()).count() * 1e-6;
}

}  // namespace

std::string KernelReport::to_string() const
{
    std::ostringstream o;
    o.setf(std::ios::fixed);
    o.precision(6);
    const double t = migration_s > 0 ? migration_s : modeling_s;
    o << "{\n  \"modeling_time\": " << modeling_s << ",\n  \"migration_time\": " << migration_s
      << ",\n  \"forward_time\": " << forward_s << ",\n  \"backward_time\": " << backward_s
      << ",\n  \"modeling_avg_time\": " << (modeling_n ? modeling_s / modeling_n : 0.0)
      << ",\n  \"migration_avg_time\": " << (migration_n ? migration_s / migration_n : 0.0)
      << ",\n  \"cell_updates\": " << cell_updates
      << ",\n  \"mcell_updates_per_s\": " << (t > 0 ? cell_updates / t / 1e6 : 0.0) << "\n}";
    return o.str();
}

RTMKernel::RTMKernel(const RTMParam &p, RTMHIPPlatform &plat) : p_(p), plat_(plat) {}

// Options beyond the reference's JSON, from the environment (the JSON schema
// and the command line stay the reference's):
//   RTM3D_WAVE_FORMAT   f32 (default) | f16 | bf16: wavefield and UPB storage
//                       (FP32 arithmetic; order 8 only; no x-section snapshots)
//   RTM3D_BACKWARD_MODE 1 (default): fused receiver step; 2: source and
//                       receiver steps in one pass; 0: unfused operators
void RTMKernel::read_options()
{
    if (const char *f = std::getenv("RTM3D_WAVE_FORMAT")) {
        if (!std::strcmp(f, "f16")) wave_format_ = RTMHIP_WAVE_F16;
        else if (!std::strcmp(f, "bf16")) wave_format_ = RTMHIP_WAVE_BF16;
        else if (!std::strcmp(f, "f32")) wave_format_ = RTMHIP_WAVE_F32;
        else throw std::runtime_error(std::string("RTM3D_WAVE_FORMAT must be f32, f16 or bf16, not ") + f);
        if (wave_format_ != RTMHIP_WAVE_F32 && p_.stencil_order != 8)
            throw std::runtime_error("RTM3D_WAVE_FORMAT f16/bf16 needs stencil_order 8");
        if (wave_format_ != RTMHIP_WAVE_F32 && p_.save_snapshots)
            throw std::runtime_error("x-section snapshots are written for FP32 wavefields only");
    }
    if (const char *m = std::getenv("RTM3D_BACKWARD_MODE")) {
        backward_mode_ = std::atoi(m);
        if (backward_mode_ < 0 || backward_mode_ > 2)
            throw std::runtime_error("RTM3D_BACKWARD_MODE must be 0, 1 or 2");
    }
}

RTMKernel::~RTMKernel() = default;

void RTMKernel::initKernel()
{
    d_.blen = (int)p_.blen;
    d_.nxe = (int)p_.nx + 2 * d_.blen;
    d_.nye = (int)p_.ny + 2 * d_.blen;
    d_.nze = (int)p_.nz + 2 * d_.blen;
    d_.order = (int)p_.stencil_order;
    hbc_ = parse_bc(p_.boundary_condition) != BoundaryCondition::RBC;
    read_options();
    const size_t E = (size_t)d_.nxe * d_.nye * d_.nze;
    // wavefield / UPB buffers in float slots: 16-bit elements take half
    auto wave_slots = [&](size_t n) { return wave_format_ == RTMHIP_WAVE_F32 ? n : (n + 1) / 2; };
    coefs_ = stencil_coefs(d_.order, p_.dx, p_.dy, p_.dz);
    const std::vector<float> g = taper_table(d_.blen, p_.taper_factor);
    taper_.createDeviceBuffer(g.size() ? g.size() : 1);
    if (!g.empty()) taper_.moveToDevice(g.data());
    const int nsrc = p_.modeling ? 2 : 4;
    for (int i = 0; i < nsrc; ++i) f_[i].createDeviceBuffer(wave_slots(E));
    const size_t nrcv = (size_t)p_.receiver_count_x * p_.receiver_count_y * p_.ntstep;
    rcv_.createDeviceBuffer(nrcv ? nrcv : 1);
    host_rcv_.assign(nrcv ? nrcv : 1, 0.f);
    if (!p_.modeling) {
        img_.createDeviceBuffer(E);
        if (hbc_)
            upb_.createDeviceBuffer(wave_slots((size_t)p_.nt * d_.nxe * d_.nye * (d_.order / 2)));
    }
}

rtmhip_rcv_geom RTMKernel::geom() const
{
    rtmhip_rcv_geom g;
    g.start_x = (int)p_.receiver_start_x;
    g.start_y = (int)p_.receiver_start_y;
    g.depth_z = (int)p_.receiver_depth_z;
    g.dist_x = (int)p_.receiver_distance_x;
    g.dist_y = (int)p_.receiver_distance_y;
    g.count_x = (int)p_.receiver_count_x;
    g.count_y = (int)p_.receiver_count_y;
    return g;
}

rtmhip_shot RTMKernel::make_shot(const ShotPos &s) const
{
    rtmhip_shot sh{};
    sh.wave_format = wave_format_;
    sh.nxe = d_.nxe;
    sh.nye = d_.nye;
    sh.nze = d_.nze;
    sh.blen = d_.blen;
    sh.order = d_.order;
    sh.nt = (int)p_.nt;
    sh.ntstep = (int)p_.ntstep;
    sh.sxe = s.sx + d_.blen;
    sh.sye = s.sy + d_.blen;
    sh.sze = s.sz + d_.blen;
    sh.rcv = geom();
    return sh;
}

// Debug x-sections (RTMFiniteDifferencesKernel.cpp:202-212): the [nye][nze]
// plane at x = sx + blen, appended to a file.
void RTMKernel::snapshot(const float *field, int sx, const std::string &file)
{
    const size_t plane = (size_t)d_.nye * d_.nze;
    std::vector<float> h(plane);
    hipcheck(rtmhip_copy_from_device(h.data(), field + (size_t)(sx + d_.blen) * plane,
                                     plane * sizeof(float)),
             "snapshot copy");
    append_raw(file, h.data(), plane);
}

void RTMKernel::progress(const char *tag, const ShotPos &s, int it, double t0)
{
    if (!p_.verbose) return;
    plat_.synchronize();
    std::printf("\r>\t%s(%d,%d,%d) = %d/%d (%.1f%%) (%.2f s)   ", tag, s.sx, s.sy, s.sz, it,
                (int)p_.nt, 100.0 * it / p_.nt, now_s() - t0);
    if (it == (int)p_.nt || it == 0) std::printf("\n");
    std::fflush(stdout);
}

void RTMKernel::rtmModel(const ShotPos &s, const DeviceGrid &v2dt2)
{
    const double t0 = now_s();
    const rtmhip_shot sh = make_shot(s);
    const std::vector<float> src = ricker((int)p_.nt, p_.dt, p_.fpeak);
    ricker_.createDeviceBuffer(src.size());
    ricker_.moveToDevice(src.data());
    f_[0].devMemSet(0.f);
    f_[1].devMemSet(0.f);
    rcv_.devMemSet(0.f);
    float *P = f_[0].ptr(), *PP = f_[1].ptr();
    const int nt = (int)p_.nt, ntstep = (int)p_.ntstep;
    std::string snapfile;
    if (p_.save_snapshots) {
        snapfile = p_.tmpdir + "/RTMSNAPSHOT_S" + std::to_string(s.sx + d_.blen) + "x" +
                   std::to_string(s.sy + d_.blen) + "_" +
                   std::to_string(nt / (int)p_.snapshot_step) + "x" + std::to_string(d_.nye) +
                   "x" + std::to_string(d_.nze) + ".snapshot.RTM_MOD";
        std::ofstream(snapfile, std::ios::out | std::ios::binary);
    }
    for (int itstep = 0; itstep < nt; itstep += ntstep) {
        int it = itstep;
        while (it < itstep + ntstep) {
            int n = itstep + ntstep - it;
            if (p_.save_snapshots) {
                // stop right after the next multiple of snapshot_step
                const int k = (int)p_.snapshot_step;
                const int next = ((it + k - 1) / k) * k;
                if (next < it + n) n = next - it + 1;
            }
            hipcheck(rtmhip_model_steps(&sh, v2dt2.ptr(), coefs_.data(), taper_.ptr(),
                                        ricker_.ptr(), &P, &PP, rcv_.ptr(), it, n),
                     "modeling steps");
            it += n;
            if (p_.save_snapshots && (it - 1) % (int)p_.snapshot_step == 0)
                snapshot(PP, s.sx, snapfile);
        }
        plat_.synchronize();
        rcv_.moveFromDevice(host_rcv_.data());
        save_raw(seismogram_name(p_.datdir, s.sx, s.sy, s.sz, (int)p_.receiver_count_x,
                                 (int)p_.receiver_count_y, itstep, ntstep),
                 host_rcv_.data(), (size_t)p_.receiver_count_x * p_.receiver_count_y * ntstep);
        progress("+ RTM_MOD", s, itstep + ntstep, t0);
    }
    plat_.synchronize();
    const double cells = (double)(d_.nxe - d_.order) * (d_.nye - d_.order) * (d_.nze - d_.order);
    report_.cell_updates += cells * nt;
    report_.modeling_s += now_s() - t0;
    report_.modeling_n++;
}

void RTMKernel::rtmMigrate(const ShotPos &s, const DeviceGrid &v2dt2)
{
    const double t0 = now_s();
    const rtmhip_shot sh = make_shot(s);
    const std::vector<float> src = ricker((int)p_.nt, p_.dt, p_.fpeak);
    ricker_.createDeviceBuffer(src.size());
    ricker_.moveToDevice(src.data());
    for (auto &f : f_) f.devMemSet(0.f);
    img_.devMemSet(0.f);
    float *P = f_[0].ptr(), *PP = f_[1].ptr();
    const int nt = (int)p_.nt, ntstep = (int)p_.ntstep;
    const char *fwdtag = hbc_ ? "HBC_FWD" : "RBC_FWD";
    const char *bwdtag = hbc_ ? "HBC_BWD" : "RBC_BWD";
    auto snapname = [&](const char *tag) {
        return p_.tmpdir + "/RTMSNAPSHOT_S" + std::to_string(s.sx) + "x" + std::to_string(s.sy) +
               "_" + std::to_string(nt / (int)p_.snapshot_step) + "x" + std::to_string(d_.nye) +
               "x" + std::to_string(d_.nze) + ".snapshot." + tag;
    };
    // ---- forward ----
    std::string snapfile;
    if (p_.save_snapshots) {
        snapfile = snapname(fwdtag);
        std::ofstream(snapfile, std::ios::out | std::ios::binary);
    }
    for (int itstep = 0; itstep < nt; itstep += ntstep) {
        int it = itstep;
        while (it < itstep + ntstep) {
            int n = itstep + ntstep - it;
            if (p_.save_snapshots) {
                const int k = (int)p_.snapshot_step;
                const int next = ((it + k - 1) / k) * k;
                if (next < it + n) n = next - it + 1;
            }
            hipcheck(rtmhip_forward_steps(&sh, hbc_ ? 1 : 0, v2dt2.ptr(), coefs_.data(),
                                          taper_.ptr(), ricker_.ptr(), &P, &PP,
                                          hbc_ ? upb_.ptr() : nullptr, it, n),
                     "forward steps");
            it += n;
            if (p_.save_snapshots && (it - 1) % (int)p_.snapshot_step == 0)
                snapshot(PP, s.sx, snapfile);
        }
        progress((std::string("+[P0] ") + fwdtag).c_str(), s, itstep + ntstep, t0);
    }
    plat_.synchronize();
    const double t1 = now_s();
    report_.forward_s += t1 - t0;
    // ---- backward ----
    if (p_.save_snapshots) {
        snapfile = snapname(bwdtag);
        std::ofstream(snapfile, std::ios::out | std::ios::binary);
    }
    float *Ps = P, *PPs = PP, *Pr = f_[2].ptr(), *PPr = f_[3].ptr();
    for (int itstep = nt - ntstep; itstep >= 0; itstep -= ntstep) {
        load_raw(seismogram_name(p_.datdir, s.sx, s.sy, s.sz, (int)p_.receiver_count_x,
                                 (int)p_.receiver_count_y, itstep, ntstep),
                 host_rcv_.data(), host_rcv_.size());
        hipcheck(rtmhip_copy_to_device(rcv_.ptr(), host_rcv_.data(),
                                       host_rcv_.size() * sizeof(float)),
                 "receiver chunk to device");
        int it = itstep + ntstep - 1;
        while (it >= itstep) {
            int n = it - itstep + 1;
            if (p_.save_snapshots) {
                const int k = (int)p_.snapshot_step;
                const int prev = (it / k) * k;  // next snapshot time going down
                if (prev >= itstep) n = it - prev + 1;
            }
            hipcheck(rtmhip_backward_steps(&sh, hbc_ ? 1 : 0, v2dt2.ptr(), coefs_.data(),
                                           taper_.ptr(), &Ps, &PPs, &Pr, &PPr,
                                           hbc_ ? upb_.ptr() : nullptr, rcv_.ptr(), img_.ptr(),
                                           it, n, backward_mode_),
                     "backward steps");
            it -= n;
            if (p_.save_snapshots && (it + 1) % (int)p_.snapshot_step == 0)
                snapshot(Pr, s.sx, snapfile);  // after the swap Pr holds the new field
        }
        progress((std::string("+[P0] ") + bwdtag).c_str(), s, nt - itstep, t1);
    }
    plat_.synchronize();
    const double t2 = now_s();
    report_.backward_s += t2 - t1;
    report_.migration_s += t2 - t0;
    report_.migration_n++;
    const double cells = (double)(d_.nxe - d_.order) * (d_.nye - d_.order) * (d_.nze - d_.order);
    report_.cell_updates += cells * (3.0 * nt - 2.0);
}

}  // namespace rtm3d
