// grid.cpp -- host-side set-up of the RTM path (see grid.hpp for sources).
#include "grid.hpp"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <stdexcept>
#include <string>

#include "param.hpp"

namespace rtm3d {

BoundaryCondition parse_bc(const std::string &s)
{
    if (s == "rbc") return BoundaryCondition::RBC;
    if (s == "abc") return BoundaryCondition::ABC;
    return BoundaryCondition::HBC;
}

void load_raw(const std::string &path, float *dst, size_t count)
{
    std::ifstream f(path, std::ios::in | std::ios::binary);
    if (!f) throw GridException("Input grid file '" + path + "' doesn't exist! Please check.");
    f.read(reinterpret_cast<char *>(dst), (std::streamsize)(count * sizeof(float)));
}

void save_raw(const std::string &path, const float *src, size_t count)
{
    std::ofstream f(path, std::ios::out | std::ios::binary);
    if (!f) throw GridException("Output grid file '" + path + "' doesn't exist! Please check.");
    f.write(reinterpret_cast<const char *>(src), (std::streamsize)(count * sizeof(float)));
}

void append_raw(const std::string &path, const float *src, size_t count)
{
    std::ofstream f(path, std::ios::out | std::ios::binary | std::ios::app);
    if (!f) throw GridException("Output grid file '" + path + "' doesn't exist! Please check.");
    f.write(reinterpret_cast<const char *>(src), (std::streamsize)(count * sizeof(float)));
}

Cube extend_borders(const Cube &vel, int b)
{
    Cube e(vel.nx + 2 * b, vel.ny + 2 * b, vel.nz + 2 * b);
    for (int x = 0; x < vel.nx; ++x)
        for (int y = 0; y < vel.ny; ++y)
            std::memcpy(&e.at(x + b, y + b, b), &vel.v[((size_t)x * vel.ny + y) * vel.nz],
                        sizeof(float) * (size_t)vel.nz);
    return e;
}

// Clamp every index into the interior (RTMGrid.cpp:371-434).
void borders_abc(Cube &e, int b)
{
    const int lx = e.nx - 2 * b, ly = e.ny - 2 * b, lz = e.nz - 2 * b;
    auto clampi = [b](int i, int l) { return i <= b ? b : (i >= l + b ? l + b - 1 : i); };
    for (int x = 0; x < e.nx; ++x) {
        const int kx = clampi(x, lx);
        for (int y = 0; y < e.ny; ++y) {
            const int ky = clampi(y, ly);
            for (int z = 0; z < e.nz; ++z) e.at(x, y, z) = e.at(kx, ky, clampi(z, lz));
        }
    }
}

namespace {

// The sequence of libc rand() after srand(seed) (glibc's default additive
// feedback generator, "TYPE_3": 31 words seeded by the Park-Miller minimal
// standard generator, 310 outputs discarded, then o = (r[n-31] + r[n-3]) >> 1
// modulo 2^32), with its own state: the reference's border fill calls rand()
// and so depends on this exact sequence, while a private state lets several
// grids be filled at once in one process.
class LibcRand {
  public:
    explicit LibcRand(unsigned seed)
    {
        int32_t w = seed ? (int32_t)seed : 1;
        ring_[0] = w;
        for (int i = 1; i < 31; ++i) {
            // 16807 * w mod (2^31 - 1) without overflow (Schrage)
            const int32_t q = w / 127773, r = w % 127773;
            w = 16807 * r - 2836 * q;
            if (w < 0) w += 2147483647;
            ring_[i] = w;
        }
        for (int i = 0; i < 310; ++i) next();
    }
    int next()
    {
        // ring_[k] holds r[n-31+k mod 31]; the newest value replaces r[n-31]
        const uint32_t v = (uint32_t)ring_[head_] + (uint32_t)ring_[(head_ + 28) % 31];
        ring_[head_] = (int32_t)v;
        head_ = (head_ + 1) % 31;
        return (int)(v >> 1);
    }

  private:
    int32_t ring_[31];
    int head_ = 3;  // the front word; the rear one is three behind
};

// randlinearval (rtm-3d/src/misc/Misc.cpp:91-102) on the libc rand() sequence.
float rand_linear(LibcRand &rng, float v, int k, int nkb)
{
    float vrand = v;
    const float dist = (float)(k * 1.0);
    const float range = (float)(nkb * 1.0);
    const float perc = (dist / range) * 100;
    const float seed = (float)((rng.next() % 100) * 1.0);
    if (seed < perc && perc > 70.)
        vrand = (float)((double)v * (1.0 / (double)perc) + (double)v * ((double)seed / 100.0));
    return vrand;
}

}  // namespace

// Random borders (RTMGrid.cpp:437-635).  The reference draws rand() inside
// OpenMP loops and is reproducible only with one thread; this fill always
// runs in that single-thread order so that its output equals the
// reference's nthreads = 1 output.
void borders_rbc(Cube &e, int b)
{
    LibcRand rng(1);  // srand(1): the sequence of a fresh process
    auto randlin = [&rng](float v, int k, int nkb) { return rand_linear(rng, v, k, nkb); };
    const int nx = e.nx, ny = e.ny, nz = e.nz;
    for (int ix = b; ix < nx - b; ix++)
        for (int iy = b; iy < ny - b; iy++)
            for (int iz = 0; iz < b; iz++) {
                const int dist = b - iz;
                e.at(ix, iy, iz) = randlin(e.at(ix, iy, b), dist, b);
                e.at(ix, iy, nz - 1 - iz) = randlin(e.at(ix, iy, nz - b - 1), dist, b);
            }
    for (int ix = b; ix < nx - b; ix++)
        for (int iz = b; iz < nz - b; iz++)
            for (int iy = 0; iy < b; iy++) {
                e.at(ix, iy, iz) = randlin(e.at(ix, b, iz), b - iy, b);
                e.at(ix, ny - b + iy, iz) = randlin(e.at(ix, ny - b - 1, iz), iy, b);
            }
    for (int iy = b; iy < ny - b; iy++)
        for (int iz = b; iz < nz - b; iz++)
            for (int ix = 0; ix < b; ix++) {
                e.at(ix, iy, iz) = randlin(e.at(b, iy, iz), b - ix, b);
                e.at(nx - b + ix, iy, iz) = randlin(e.at(nx - b - 1, iy, iz), ix, b);
            }
    // x-z corner columns over all y: y factor (if any), then x, then z.
    for (int ix = 0; ix < b; ix++)
        for (int iz = 0; iz < b; iz++)
            for (int iy = 0; iy < ny; iy++) {
                const int ky = iy < b ? b : (iy >= ny - b ? ny - b - 1 : iy);
                struct Corner {
                    int sx, sz, ox, oz, kx, kz;
                } cs[4] = {
                    {b, b, ix, iz, b - ix, b - iz},
                    {nx - b - 1, b, nx - b + ix, iz, ix, b - iz},
                    {b, nz - b - 1, ix, nz - b + iz, b - ix, iz},
                    {nx - b - 1, nz - b - 1, nx - b + ix, nz - b + iz, ix, iz},
                };
                for (const Corner &c : cs) {
                    float v = e.at(c.sx, ky, c.sz);
                    if (iy < b) v = randlin(v, b - iy, b);
                    else if (iy >= ny - b) v = randlin(v, iy - (ny - b), b);
                    v = randlin(v, c.kx, b);
                    v = randlin(v, c.kz, b);
                    e.at(c.ox, iy, c.oz) = v;
                }
            }
    for (int iy = 0; iy < b; iy++)
        for (int iz = 0; iz < b; iz++)
            for (int ix = b; ix < nx - b; ix++) {
                float v;
                v = randlin(e.at(ix, b, b), b - iy, b);
                e.at(ix, iy, iz) = randlin(v, b - iz, b);
                v = randlin(e.at(ix, ny - b - 1, b), iy, b);
                e.at(ix, ny - b + iy, iz) = randlin(v, b - iz, b);
                v = randlin(e.at(ix, b, nz - b - 1), b - iy, b);
                e.at(ix, iy, nz - b + iz) = randlin(v, iz, b);
                v = randlin(e.at(ix, ny - b - 1, nz - b - 1), iy, b);
                e.at(ix, ny - b + iy, nz - b + iz) = randlin(v, iz, b);
            }
    for (int ix = 0; ix < b; ix++)
        for (int iy = 0; iy < b; iy++)
            for (int iz = b; iz < nz - b; iz++) {
                float v;
                v = randlin(e.at(b, b, iz), b - ix, b);
                e.at(ix, iy, iz) = randlin(v, b - iy, b);
                v = randlin(e.at(b, ny - b - 1, iz), b - ix, b);
                e.at(ix, ny - b + iy, iz) = randlin(v, iy, b);
                v = randlin(e.at(nx - b - 1, b, iz), ix, b);
                e.at(nx - b + ix, iy, iz) = randlin(v, b - iy, b);
                v = randlin(e.at(nx - b - 1, ny - b - 1, iz), ix, b);
                e.at(nx - b + ix, ny - b + iy, iz) = randlin(v, iy, b);
            }
}

// ABC, then RBC, then the top slab copies z = blen (RTMGrid.cpp:638-696).
void borders_hbc(Cube &e, int b)
{
    borders_abc(e, b);
    borders_rbc(e, b);
    const int lx = e.nx - 2 * b, ly = e.ny - 2 * b;
    auto clampi = [b](int i, int l) { return i <= b ? b : (i >= l + b ? l + b - 1 : i); };
    for (int x = 0; x < e.nx; ++x)
        for (int y = 0; y < e.ny; ++y) {
            const float v = e.at(clampi(x, lx), clampi(y, ly), b);
            for (int z = 0; z < b; ++z) e.at(x, y, z) = v;
        }
}

void apply_v2dt2(Cube &e, float dt)
{
    const float dt2 = dt * dt;
    for (float &v : e.v) {
        v *= v;
        v *= dt2;
    }
}

std::vector<float> laplacian_taps(int order)
{
    std::vector<float> c((size_t)order + 1, 0.f);
    switch (order) {
    case 2: c = {1.f, -2.f, 1.f}; break;
    case 4:
        c = {(float)(-1. / 12.), (float)(4. / 3.), (float)(-5. / 2.), (float)(4. / 3.),
             (float)(-1. / 12.)};
        break;
    case 6:
        c = {(float)(1. / 90.), (float)(-3. / 20.), (float)(3. / 2.), (float)(-49. / 18.),
             (float)(3. / 2.), (float)(-3. / 20.), (float)(1. / 90.)};
        break;
    case 8:
        c = {(float)(-1. / 560.), (float)(8. / 315.), (float)(-1. / 5.), (float)(8. / 5.),
             (float)(-205. / 72.), (float)(8. / 5.), (float)(-1. / 5.), (float)(8. / 315.),
             (float)(-1. / 560.)};
        break;
    default:
        // the reference derives other even orders from a windowed series
        // (Misc.cpp:106-137); the device kernels run orders 2 to 8 only
        throw std::runtime_error("stencil order " + std::to_string(order) +
                                 " not supported (2, 4, 6 or 8)");
    }
    return c;
}

std::vector<float> stencil_coefs(int order, float dx, float dy, float dz)
{
    const std::vector<float> taps = laplacian_taps(order);
    const float d[3] = {dx, dy, dz};
    std::vector<float> out(3 * ((size_t)order + 1));
    for (int k = 0; k < 3; ++k) {
        const float devinv = (float)((1.0 / (double)d[k]) * (1.0 / (double)d[k]));
        for (int io = 0; io <= order; ++io) out[(size_t)k * (order + 1) + io] = devinv * taps[io];
    }
    return out;
}

std::vector<float> ricker(int nt, float dt, float fpeak)
{
    std::vector<float> s((size_t)nt);
    for (int it = 0; it < nt; ++it) {
        const float t = (float)((double)((float)it * dt) - 1.0 / (double)fpeak);
        const float x = (float)(M_PI * (double)fpeak * (double)t);
        const float xx = x * x;
        s[it] = (float)((double)std::exp(-xx) * (1.0 - 2.0 * (double)xx));
    }
    return s;
}

std::vector<float> taper_table(int blen, float factor)
{
    std::vector<float> g((size_t)std::max(blen, 0));
    const float dfrac = (float)((double)std::sqrt(-std::log(factor)) / (1. * blen));
    for (int i = 0; i < blen; ++i)
        g[i] = (float)std::exp(-std::pow((double)(dfrac * (float)i), 2.0));
    return g;
}

std::string seismogram_name(const std::string &datdir, int sx, int sy, int sz, int nrx, int nry,
                            int itstep, int ntstep)
{
    return datdir + "/RTMSEISMOGRAM_S" + std::to_string(sx) + "x" + std::to_string(sy) + "x" +
           std::to_string(sz) + "_NXNY_" + std::to_string(nrx) + "x" + std::to_string(nry) +
           "_NTSTART" + std::to_string(itstep) + "_NTSTEP" + std::to_string(ntstep) + ".seism";
}

std::string image_name(const std::string &outdir, std::string mname, int ns, int nx, int ny,
                       int nz, int nt, int np)
{
    std::transform(mname.begin(), mname.end(), mname.begin(),
                   [](unsigned char c) { return (char)std::toupper(c); });
    return outdir + "/RTMIMG_" + mname + "_NS" + std::to_string(ns) + "_NXNYNZ_" +
           std::to_string(nx) + "x" + std::to_string(ny) + "x" + std::to_string(nz) + "_NT" +
           std::to_string(nt) + ".bin.NP" + std::to_string(np);
}

// RTMFreqDomainKernel::initKernel, rtmInitWList and rtmInitFreqKernels
// (rtm-3d/src/kernel/RTMFiniteDifferencesKernel_FreqDomain.cpp:17-52,
// 125-152) with the reference's float members and promotions: df is the
// float of 1/(double(nt)*dt); nw truncates the float (fmax - fmin)/df; the
// angular step w = float(2*double(pi_f)*(float(iw)*df + fmin)*double(dt));
// the kernels are cosf / sinf of float(double(w)*it); w2 = w*w in float.
FreqTables freq_tables(int nt, float dt, float fmin, float fmax)
{
    FreqTables t;
    const float freq_max = 1.0 * fmax;
    const float freq_min = 1.0 * fmin;
    const float df = 1. / ((nt * 1.0) * dt);
    const float nf = ((freq_max - freq_min) / df);
    t.nw = nf > 0 ? (int)(size_t)nf : 0;
    const float pi = acosf(-1.);
    t.w2.resize(t.nw);
    t.kre.resize((size_t)t.nw * nt);
    t.kim.resize((size_t)t.nw * nt);
    for (int iw = 0; iw < t.nw; ++iw) {
        const float lw = (float)iw * 1.0;
        const float w = (float)(2. * pi * ((lw * df) + freq_min) * dt);
        for (int it = 0; it < nt; ++it) {
            t.kre[(size_t)iw * nt + it] = cosf(w * (it * 1.0));
            t.kim[(size_t)iw * nt + it] = sinf(w * (it * 1.0));
        }
        t.w2[iw] = w * w;
    }
    return t;
}

}  // namespace rtm3d
