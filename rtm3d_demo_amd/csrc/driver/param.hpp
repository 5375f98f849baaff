// param.hpp -- RTM input parameters read from the reference's JSON schema
// (rtm-3d/include/RTMParam.hpp:117-283, rtm-3d/src/param/RTMParam.cpp:99-264).
#pragma once
#include <stdexcept>
#include <string>

namespace rtm3d {

// Error classes mirroring RTMException / RTMParamException / GridException
// (rtm-3d/include/RTMException.hpp): all end the program with exit status 1.
struct RTMException : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct RTMParamException : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct GridException : std::runtime_error {
    using std::runtime_error::runtime_error;
};

struct RTMParam {
    std::string mname, vpfile, vpefile, outdir, tmpdir, rtm_type, datdir, boundary_condition;
    std::string fpga_fwd_kernel, fpga_bwd_kernel, fpga_xclbin;
    unsigned nx = 0, ny = 0, nz = 0, blen = 0, width_m = 0, height_m = 0, depth_m = 0, nt = 0;
    unsigned stencil_order = 8;
    unsigned source_total = 0, source_start_x = 0, source_start_y = 0, source_depth_z = 0;
    unsigned source_distance_x = 0, source_distance_y = 0, source_count_x = 0,
             source_count_y = 0;
    unsigned receiver_total = 0, receiver_start_x = 0, receiver_start_y = 0,
             receiver_depth_z = 0;
    unsigned receiver_distance_x = 0, receiver_distance_y = 0, receiver_count_x = 0,
             receiver_count_y = 0;
    unsigned ntstep = 0, snapshot_step = 1, nthreads = 1;
    float dz = 0, dy = 0, dx = 0, dt = 0, fpeak = 0, fmax = 0, taper_factor = 0;
    bool modeling = false, migration = false, save_vpe_file = false, load_vpe_from_file = false;
    bool verbose = false, save_snapshots = false, distributed_grid = false;
    bool fmig = false, fmig_distributed_imaging = false;
    float fmig_max_freq = 0, fmig_min_freq = 0;
    unsigned fmig_nwstep = 1;

    // Parses and checks the file; throws RTMParamException on bad input.
    void load(const std::string &json_path);
    std::string to_string() const;
};

}  // namespace rtm3d
