// param.cpp -- loads RTMParam from JSON (rtm-3d/src/param/RTMParam.cpp:99-264).
//
// Every key the reference reads is required (the reference fails on a
// missing key because reading it yields null); numbers may be written as
// integers or decimals; booleans and strings must have their JSON types.
// As in the reference, nt < ntstep is rejected (RTMParam.cpp:241-245).
#include "param.hpp"

#include <sstream>

#include "json.hpp"

namespace rtm3d {

namespace {

std::string str(const Json &j, const char *k)
{
    try {
        return j.at(k).as_string();
    } catch (const JsonError &e) {
        throw RTMParamException(std::string("[ Input parameter '") + k + "': " + e.what() + " ]");
    }
}

bool boolean(const Json &j, const char *k)
{
    try {
        return j.at(k).as_bool();
    } catch (const JsonError &e) {
        throw RTMParamException(std::string("[ Input parameter '") + k + "': " + e.what() + " ]");
    }
}

unsigned uint(const Json &j, const char *k)
{
    try {
        const long long v = j.at(k).as_int();
        if (v < 0) throw JsonError("must not be negative");
        return (unsigned)v;
    } catch (const JsonError &e) {
        throw RTMParamException(std::string("[ Input parameter '") + k + "': " + e.what() + " ]");
    }
}

float real(const Json &j, const char *k)
{
    try {
        return (float)j.at(k).as_double();
    } catch (const JsonError &e) {
        throw RTMParamException(std::string("[ Input parameter '") + k + "': " + e.what() + " ]");
    }
}

}  // namespace

void RTMParam::load(const std::string &path)
{
    Json j;
    try {
        j = Json::parse_file(path);
    } catch (const JsonError &e) {
        throw RTMParamException(e.what());
    }
    mname = str(j, "mname");
    vpfile = str(j, "vpfile");
    vpefile = str(j, "vpefile");
    outdir = str(j, "outdir");
    tmpdir = str(j, "tmpdir");
    rtm_type = str(j, "rtm_type");
    boundary_condition = str(j, "boundary_condition");
    datdir = str(j, "datdir");
    nx = uint(j, "nx");
    ny = uint(j, "ny");
    nz = uint(j, "nz");
    blen = uint(j, "blen");
    width_m = uint(j, "width_m");
    height_m = uint(j, "height_m");
    depth_m = uint(j, "depth_m");
    nt = uint(j, "nt");
    stencil_order = uint(j, "stencil_order");
    source_total = uint(j, "source_total");
    source_start_x = uint(j, "source_start_x");
    source_start_y = uint(j, "source_start_y");
    source_depth_z = uint(j, "source_depth_z");
    source_distance_x = uint(j, "source_distance_x");
    source_distance_y = uint(j, "source_distance_y");
    source_count_x = uint(j, "source_count_x");
    source_count_y = uint(j, "source_count_y");
    receiver_total = uint(j, "receiver_total");
    receiver_start_x = uint(j, "receiver_start_x");
    receiver_start_y = uint(j, "receiver_start_y");
    receiver_depth_z = uint(j, "receiver_depth_z");
    receiver_distance_x = uint(j, "receiver_distance_x");
    receiver_distance_y = uint(j, "receiver_distance_y");
    receiver_count_x = uint(j, "receiver_count_x");
    receiver_count_y = uint(j, "receiver_count_y");
    ntstep = uint(j, "ntstep");
    nthreads = uint(j, "nthreads");
    dz = real(j, "dz");
    dy = real(j, "dy");
    dx = real(j, "dx");
    dt = real(j, "dt");
    fpeak = real(j, "fpeak");
    fmax = real(j, "fmax");
    taper_factor = real(j, "taper_factor");
    modeling = boolean(j, "modeling");
    migration = boolean(j, "migration");
    save_vpe_file = boolean(j, "save_vpe_file");
    load_vpe_from_file = boolean(j, "load_vpe_from_file");
    verbose = boolean(j, "verbose");
    save_snapshots = boolean(j, "save_snapshots");
    snapshot_step = uint(j, "snapshot_step");
    distributed_grid = boolean(j, "distributed_grid");
    fmig = boolean(j, "fmig");
    fmig_max_freq = real(j, "fmig_max_freq");
    fmig_min_freq = real(j, "fmig_min_freq");
    fmig_nwstep = uint(j, "fmig_nwstep");
    fmig_distributed_imaging = boolean(j, "fmig_distributed_imaging");
    fpga_xclbin = str(j, "fpga_xclbin");
    fpga_fwd_kernel = str(j, "fpga_fwd_kernel");
    fpga_bwd_kernel = str(j, "fpga_bwd_kernel");

    if (nt < ntstep) throw RTMParamException("[ RTMParam Error: NT smaller than NTSTEP ]");
    if (fmig && fmig_max_freq <= fmig_min_freq)
        throw RTMParamException("[ RTMParam Error: freq_max smaller than freq_min ]");
    if (fmig && distributed_grid)
        throw RTMParamException(
            "[ RTMParam Error: Can't use 'distributed_grid' on fmig migrations. Set it to false "
            "on input json file ]");
}

std::string RTMParam::to_string() const
{
    std::ostringstream o;
    o << "#####################################################################";
    o << "\n  # mname              = " << mname << "\n  # vpfile             = " << vpfile
      << "\n  # vpefile            = " << vpefile << "\n  # outdir             = " << outdir
      << "\n  # tmpdir             = " << tmpdir << "\n  # datdir             = " << datdir
      << "\n  # boundary_condition = " << boundary_condition << "\n  # nx, ny, nz         = "
      << nx << ", " << ny << ", " << nz << "\n  # blen               = " << blen
      << "\n  # nt, ntstep         = " << nt << ", " << ntstep
      << "\n  # dx, dy, dz, dt     = " << dx << ", " << dy << ", " << dz << ", " << dt
      << "\n  # fpeak              = " << fpeak << "\n  # taper_factor       = " << taper_factor
      << "\n  # stencil_order      = " << stencil_order << "\n  # sources            = "
      << source_count_x << "x" << source_count_y << " from (" << source_start_x << ","
      << source_start_y << "," << source_depth_z << ") step (" << source_distance_x << ","
      << source_distance_y << ")"
      << "\n  # receivers          = " << receiver_count_x << "x" << receiver_count_y
      << " from (" << receiver_start_x << "," << receiver_start_y << "," << receiver_depth_z
      << ") step (" << receiver_distance_x << "," << receiver_distance_y << ")"
      << "\n  # modeling/migration = " << modeling << "/" << migration
      << "\n  # save/load vpe      = " << save_vpe_file << "/" << load_vpe_from_file
      << "\n  # snapshots          = " << save_snapshots << " every " << snapshot_step;
    o << "\n> #####################################################################";
    return o.str();
}

}  // namespace rtm3d
