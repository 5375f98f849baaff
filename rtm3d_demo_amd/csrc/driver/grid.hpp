// grid.hpp -- host-side grids and the set-up arithmetic of the RTM path.
//
// Restates, with the reference's float/double promotions, the host
// computations whose results feed the device: velocity extension and border
// fills (rtm-3d/src/grid/RTMGrid.cpp:47-86, :371-696), v^2 dt^2
// (rtm-3d/src/controller/RTMController.cpp:110-111), stencil taps
// (rtm-3d/src/misc/Misc.cpp:106-176, rtm-3d/include/RTMGrid.hpp:433-446),
// Ricker samples (rtm-3d/include/RTMSeismic.hpp:135-198) and the Gaussian
// taper (rtm-3d/include/RTMSeismic.hpp:294-305).  Also the reference's file
// naming (rtm-3d/include/RTMBase.hpp:79-138) and raw binary IO
// (rtm-3d/include/Grid.hpp:248-311).
#pragma once
#include <cstddef>
#include <string>
#include <vector>

namespace rtm3d {

// A [nx][ny][nz] float grid, z fastest.
struct Cube {
    int nx = 0, ny = 0, nz = 0;
    std::vector<float> v;
    Cube() = default;
    Cube(int x, int y, int z) : nx(x), ny(y), nz(z), v((size_t)x * y * z, 0.f) {}
    size_t size() const { return v.size(); }
    float &at(int x, int y, int z) { return v[((size_t)x * ny + y) * nz + z]; }
    float at(int x, int y, int z) const { return v[((size_t)x * ny + y) * nz + z]; }
};

enum class BoundaryCondition { ABC, HBC, RBC };
// "hbc", "rbc", "abc"; anything else means HBC (rtm-3d/include/RTMGrid.hpp:39-57).
BoundaryCondition parse_bc(const std::string &s);

// Raw float files; throw GridException when a file cannot be opened.
void load_raw(const std::string &path, float *dst, size_t count);
void save_raw(const std::string &path, const float *src, size_t count);
void append_raw(const std::string &path, const float *src, size_t count);

Cube extend_borders(const Cube &vel, int blen);
void borders_abc(Cube &ve, int blen);
void borders_rbc(Cube &ve, int blen);  // libc rand(), reseeded with 1, one thread
void borders_hbc(Cube &ve, int blen);
void apply_v2dt2(Cube &ve, float dt);

std::vector<float> laplacian_taps(int order);
std::vector<float> stencil_coefs(int order, float dx, float dy, float dz);
std::vector<float> ricker(int nt, float dt, float fpeak);
std::vector<float> taper_table(int blen, float factor);
// Frequency-domain migration tables: the number of frequencies nw, the
// squared angular steps w2[nw] and the DFT kernels kre / kim [nw][nt].
struct FreqTables {
    int nw = 0;
    std::vector<float> w2, kre, kim;
};
FreqTables freq_tables(int nt, float dt, float fmin, float fmax);

std::string seismogram_name(const std::string &datdir, int sx, int sy, int sz, int nrx, int nry,
                            int itstep, int ntstep);
std::string image_name(const std::string &outdir, std::string mname, int ns, int nx, int ny,
                       int nz, int nt, int np);

}  // namespace rtm3d
