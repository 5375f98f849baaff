// kernel.cpp -- per-shot time loops (see kernel.hpp).
#include "kernel.hpp"

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <utility>

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
//   RTM3D_UPB_SPILL     1: keep one receiver chunk of the HBC upper-border
//                       history on the device and the rest in host memory,
//                       as the reference keeps one chunk and spills the rest
//                       to files (the default holds every step on the device
//                       and spills only when that allocation fails)
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
    if (const char *u = std::getenv("RTM3D_UPB_SPILL")) {
        if (std::strcmp(u, "0") && std::strcmp(u, "1"))
            throw std::runtime_error("RTM3D_UPB_SPILL must be 0 or 1");
        upb_spill_ = !std::strcmp(u, "1");
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
        img_.devMemSet(0.f);
        if (p_.fmig) {
            if (wave_format_ != RTMHIP_WAVE_F32)
                throw std::runtime_error("frequency-domain migration runs FP32 wavefields only");
            freq_ = freq_tables((int)p_.nt, p_.dt, p_.fmig_min_freq, p_.fmig_max_freq);
            if (freq_.nw <= 0)
                throw std::runtime_error("frequency-domain migration: no frequency in the range");
            kre_.createDeviceBuffer(freq_.kre.size());
            kre_.moveToDevice(freq_.kre.data());
            kim_.createDeviceBuffer(freq_.kim.size());
            kim_.moveToDevice(freq_.kim.data());
            w2_.createDeviceBuffer(freq_.w2.size());
            w2_.moveToDevice(freq_.w2.data());
            // every frequency in one group (the reference CPU platform's
            // choice, RTMFiniteDifferencesKernel_FreqDomain.cpp:133-142) when
            // the four accumulators can be allocated; otherwise groups of
            // fmig_nwstep, as the reference's accelerator builds do
            const size_t vol = (size_t)p_.nx * p_.ny * p_.nz;
            nwstep_ = freq_.nw;
            try {
                for (auto &a : acc_) a.createDeviceBuffer(vol * (size_t)nwstep_);
            } catch (const RTMException &) {
                for (auto &a : acc_) a.removeDeviceBuffer();
                if (p_.fmig_nwstep == 0) throw std::runtime_error("fmig_nwstep must be positive");
                nwstep_ = (int)std::min<unsigned>(p_.fmig_nwstep, (unsigned)freq_.nw);
                for (auto &a : acc_) a.createDeviceBuffer(vol * (size_t)nwstep_);
            }
        } else if (hbc_) {
            // the reference holds one receiver chunk of the history and
            // spills each chunk to a file (RTMFiniteDifferencesKernel_HBC.cpp
            // :34, :188-195, :261-277); here the whole history stays on the
            // device unless it does not fit or a spill is asked for, and the
            // spilled chunks stay in host memory
            const size_t layer = (size_t)d_.nxe * d_.nye * (d_.order / 2);
            if (!upb_spill_) {
                try {
                    upb_.createDeviceBuffer(wave_slots((size_t)p_.nt * layer));
                } catch (const RTMException &) {
                    upb_.removeDeviceBuffer();
                    upb_spill_ = true;
                }
            }
            if (upb_spill_) {
                const size_t es = wave_format_ == RTMHIP_WAVE_F32 ? 4 : 2;
                upb_chunk_bytes_ = (size_t)p_.ntstep * layer * es;
                upb_.createDeviceBuffer(wave_slots((size_t)p_.ntstep * layer));
                const size_t chunks = ((size_t)p_.nt + p_.ntstep - 1) / p_.ntstep;
                host_upb_.resize(upb_chunk_bytes_ * chunks);
            }
        }
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
    if (p_.fmig) return rtmMigrateFreq(s, v2dt2);
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
    const int upb_mode = !hbc_ ? 0 : upb_spill_ ? 2 : 1;  // history layout
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
            hipcheck(rtmhip_forward_steps(&sh, upb_mode, v2dt2.ptr(), coefs_.data(),
                                          taper_.ptr(), ricker_.ptr(), &P, &PP,
                                          hbc_ ? upb_.ptr() : nullptr, it, n),
                     "forward steps");
            it += n;
            if (p_.save_snapshots && (it - 1) % (int)p_.snapshot_step == 0)
                snapshot(PP, s.sx, snapfile);
        }
        if (upb_spill_) {
            unsigned char *slot = host_upb_.data() + upb_chunk_bytes_ * (itstep / ntstep);
            hipcheck(rtmhip_copy_from_device(slot, upb_.ptr(), upb_chunk_bytes_),
                     "upper-border chunk to host");
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
        if (upb_spill_)
            hipcheck(rtmhip_copy_to_device(upb_.ptr(),
                                           host_upb_.data() + upb_chunk_bytes_ * (itstep / ntstep),
                                           upb_chunk_bytes_),
                     "upper-border chunk to device");
        int it = itstep + ntstep - 1;
        while (it >= itstep) {
            int n = it - itstep + 1;
            if (p_.save_snapshots) {
                const int k = (int)p_.snapshot_step;
                const int prev = (it / k) * k;  // next snapshot time going down
                if (prev >= itstep) n = it - prev + 1;
            }
            hipcheck(rtmhip_backward_steps(&sh, upb_mode, v2dt2.ptr(), coefs_.data(),
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

// Frequency-domain migration of one shot
// (RTMFreqDomainKernel::rtmFreqDomainPropagation and rtmUpdatePressureGrids,
// rtm-3d/src/kernel/RTMFiniteDifferencesKernel_FreqDomain.cpp:220-369).  Per
// group of frequencies, from zero wavefields and accumulators: per receiver
// chunk (last first), steps lt descending with rt = itstep + lt and
// st = nt-1-rt: both fields stepped (rtmStepMultipleWave: two independent
// steps), all borders of the four grids tapered, the source at st added to
// PPs, the receivers of lt set into PPr, the contributions and the image of
// the group's frequencies at st, then the swaps.  The image grid is not
// cleared between shots, as in the reference: each shot image handed to the
// stack is the running sum of the shots so far.
void RTMKernel::rtmMigrateFreq(const ShotPos &s, const DeviceGrid &v2dt2)
{
    const double t0 = now_s();
    const int nt = (int)p_.nt, ntstep = (int)p_.ntstep;
    const std::vector<float> src = ricker(nt, p_.dt, p_.fpeak);
    ricker_.createDeviceBuffer(src.size());
    ricker_.moveToDevice(src.data());
    rtmhip_shot sh = make_shot(s);
    sh.wave_format = RTMHIP_WAVE_F32;  // the frequency-domain path keeps FP32 fields
    std::string snapfile;
    if (p_.save_snapshots) {
        snapfile = p_.tmpdir + "/RTMSNAPSHOT_S" + std::to_string(s.sx) + "x" +
                   std::to_string(s.sy) + "_" + std::to_string(nt / (int)p_.snapshot_step) +
                   "x" + std::to_string(d_.nye) + "x" + std::to_string(d_.nze) +
                   ".snapshot.FREQIMG";
        std::ofstream(snapfile, std::ios::out | std::ios::binary);
    }
    for (int gw = 0; gw < freq_.nw; gw += nwstep_) {
        for (auto &f : f_) f.devMemSet(0.f);
        for (auto &a : acc_) a.devMemSet(0.f);
        float *Ps = f_[0].ptr(), *PPs = f_[1].ptr(), *Pr = f_[2].ptr(), *PPr = f_[3].ptr();
        for (int itstep = nt - ntstep; itstep >= 0; itstep -= ntstep) {
            load_raw(seismogram_name(p_.datdir, s.sx, s.sy, s.sz, (int)p_.receiver_count_x,
                                     (int)p_.receiver_count_y, itstep, ntstep),
                     host_rcv_.data(), host_rcv_.size());
            hipcheck(rtmhip_copy_to_device(rcv_.ptr(), host_rcv_.data(),
                                           host_rcv_.size() * sizeof(float)),
                     "receiver chunk to device");
            // steps lt = ntstep-1 .. 0 of the chunk through the library's
            // loop (rtmhip_freq_steps: the same operator sequence per step),
            // split after each step whose image the snapshot option saves
            int lt = ntstep - 1;
            while (lt >= 0) {
                int n = lt + 1;
                if (p_.save_snapshots) {
                    for (int j = lt; j >= 0; --j)
                        if ((nt - 1 - (itstep + j)) % (int)p_.snapshot_step == 0) {
                            n = lt - j + 1;
                            break;
                        }
                }
                hipcheck(rtmhip_freq_steps(&sh, v2dt2.ptr(), coefs_.data(), taper_.ptr(),
                                           ricker_.ptr(), &Ps, &PPs, &Pr, &PPr, rcv_.ptr(),
                                           img_.ptr(), acc_[0].ptr(), acc_[1].ptr(),
                                           acc_[2].ptr(), acc_[3].ptr(), kre_.ptr(), kim_.ptr(),
                                           w2_.ptr(), nwstep_, gw, freq_.nw, itstep + lt, n),
                         "frequency-domain steps");
                lt -= n;
                const int st = nt - 1 - (itstep + lt + 1);
                if (p_.save_snapshots && st % (int)p_.snapshot_step == 0)
                    snapshot(img_.ptr(), s.sx, snapfile);
            }
            progress("+ FMIG", s, nt - itstep, t0);
        }
    }
    plat_.synchronize();
    const double t1 = now_s();
    report_.migration_s += t1 - t0;
    report_.migration_n++;
    const double cells =
        (double)(d_.nxe - d_.order) * (d_.nye - d_.order) * (d_.nze - d_.order);
    report_.cell_updates += cells * 2.0 * nt * (double)((freq_.nw + nwstep_ - 1) / nwstep_);
}

}  // namespace rtm3d
