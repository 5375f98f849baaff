// host_api.cpp -- C entry points of the host library (librtm3d_host.so) for
// scripting and tests: the set-up arithmetic the driver feeds to the device,
// and an in-process run of a JSON input.
#include <cstring>
#include <exception>
#include <string>

#include "controller.hpp"
#include "grid.hpp"

using namespace rtm3d;

extern "C" {

// Returns 0, or 1 for an order the device kernels do not run (out untouched).
int rtm3d_stencil_coefs(int order, float dx, float dy, float dz, float *out)
{
    try {
        const std::vector<float> c = stencil_coefs(order, dx, dy, dz);
        std::memcpy(out, c.data(), c.size() * sizeof(float));
        return 0;
    } catch (const std::exception &) {
        return 1;
    }
}

void rtm3d_ricker(int nt, float dt, float fpeak, float *out)
{
    const std::vector<float> s = ricker(nt, dt, fpeak);
    std::memcpy(out, s.data(), s.size() * sizeof(float));
}

void rtm3d_taper_table(int blen, float factor, float *out)
{
    const std::vector<float> g = taper_table(blen, factor);
    std::memcpy(out, g.data(), g.size() * sizeof(float));
}

// bc: 0 = ABC, 1 = HBC, 2 = RBC.  out holds (nx+2b)(ny+2b)(nz+2b) floats.
void rtm3d_extend_velocity(const float *vel, int nx, int ny, int nz, int blen, int bc,
                           float *out)
{
    Cube v(nx, ny, nz);
    std::memcpy(v.v.data(), vel, v.size() * sizeof(float));
    Cube e = extend_borders(v, blen);
    if (bc == 0) borders_abc(e, blen);
    else if (bc == 2) borders_rbc(e, blen);
    else borders_hbc(e, blen);
    std::memcpy(out, e.v.data(), e.size() * sizeof(float));
}

void rtm3d_v2dt2(float *v, size_t n, float dt)
{
    const float dt2 = dt * dt;
    for (size_t i = 0; i < n; ++i) {
        v[i] *= v[i];
        v[i] *= dt2;
    }
}

// Frequency-domain migration tables; returns nw and writes w2[nw] and
// kre / kim [nw][nt] when nw <= cap.
int rtm3d_freq_tables(int nt, float dt, float fmin, float fmax, float *w2, float *kre, float *kim,
                      int cap)
{
    const FreqTables t = freq_tables(nt, dt, fmin, fmax);
    if (t.nw <= cap && t.nw > 0) {
        std::memcpy(w2, t.w2.data(), t.w2.size() * sizeof(float));
        std::memcpy(kre, t.kre.data(), t.kre.size() * sizeof(float));
        std::memcpy(kim, t.kim.data(), t.kim.size() * sizeof(float));
    }
    return t.nw;
}

int rtm3d_run_json(const char *path, int device) { return run_json(path, device); }

}  // extern "C"
