// controller.cpp -- run driver (see controller.hpp).
#include "controller.hpp"

#include <omp.h>

#include <cstdio>
#include <exception>
#include <iostream>
#include <vector>

#include "grid.hpp"
#include "json.hpp"
#include "kernel.hpp"
#include "platform.hpp"

namespace rtm3d {

namespace {

void say(const RTMParam &p, const std::string &m)
{
    if (p.verbose) std::cout << "> " << m << std::endl << std::flush;
}

// Shot descriptors (RTMController.cpp:122-166): x outer, y inner.
std::vector<ShotPos> shot_positions(const RTMParam &p)
{
    std::vector<ShotPos> s;
    for (unsigned ix = 0; ix < p.source_count_x; ++ix)
        for (unsigned iy = 0; iy < p.source_count_y; ++iy) {
            ShotPos sp{(int)(ix * p.source_distance_x + p.source_start_x),
                       (int)(iy * p.source_distance_y + p.source_start_y), (int)p.source_depth_z};
            if (sp.sx >= (int)p.nx || sp.sy >= (int)p.ny || sp.sz >= (int)p.nz)
                throw RTMException("[ Invalid Source Position: (" + std::to_string(sp.sx) + "," +
                                   std::to_string(sp.sy) + "," + std::to_string(sp.sz) + ") ]");
            s.push_back(sp);
        }
    return s;
}

// Velocity model -> extended v^2 dt^2 (RTMController.cpp:27-119).
Cube prepare_velocity(const RTMParam &p)
{
    const int b = (int)p.blen;
    Cube vel((int)p.nx, (int)p.ny, (int)p.nz);
    say(p, "Loading Velocity Model '" + p.mname + "' at '" + p.vpfile + "'.");
    load_raw(p.vpfile, vel.v.data(), vel.size());
    say(p, "Extending '" + p.mname + "' borders (blen=" + std::to_string(b) + ")");
    Cube ve = extend_borders(vel, b);
    if (p.modeling) {
        say(p, "Initializing velocity model with 'ABC' borders... ");
        borders_abc(ve, b);
    } else if (p.load_vpe_from_file) {
        say(p, "Loading Extended Velocity Model '" + p.mname + "' from '" + p.vpefile + "'.");
        load_raw(p.vpefile, ve.v.data(), ve.size());
    } else {
        const BoundaryCondition bc = parse_bc(p.boundary_condition);
        say(p, "Initializing velocity model with '" + p.boundary_condition + "' borders... ");
        if (bc == BoundaryCondition::ABC) borders_abc(ve, b);
        else if (bc == BoundaryCondition::RBC) borders_rbc(ve, b);
        else borders_hbc(ve, b);
    }
    if (p.save_vpe_file) {
        say(p, "Saving Extended Velocity Model '" + p.mname + "' to '" + p.vpefile + "'.");
        save_raw(p.vpefile, ve.v.data(), ve.size());
    }
    apply_v2dt2(ve, p.dt);
    return ve;
}

}  // namespace

void run(RTMParam &p, int device)
{
    if (p.nx <= 1) throw RTMException("2D models (nx == 1) are not provided by this build");
    if (p.receiver_count_x > p.nx || p.receiver_count_y > p.ny)
        throw RTMException("[ Invalid Receiver Grid Size: (" + std::to_string(p.receiver_count_x) +
                           "," + std::to_string(p.receiver_count_y) + "," + std::to_string(p.nt) +
                           ") ]");
    if (p.nt % p.ntstep != 0)
        throw RTMParamException("[ RTMParam Error: NTSTEP must divide NT ]");
    if (p.save_snapshots && p.snapshot_step == 0)
        throw RTMParamException("[ RTMParam Error: snapshot_step must be positive ]");
    if (p.nthreads > 0) omp_set_num_threads((int)p.nthreads);
    say(p, "Initializing RTM Acoustic Process... ");
    say(p, p.to_string());
    if (p.distributed_grid) say(p, "distributed_grid: one device holds the whole model");

    RTMHIPPlatform plat(device);
    plat.initRTMPlatform();
    struct Finalize {
        RTMHIPPlatform &pl;
        ~Finalize() { pl.destroyRTMPlatform(); }
    } fin{plat};

    const std::vector<ShotPos> shots = shot_positions(p);
    Cube v2 = prepare_velocity(p);
    const int b = (int)p.blen;
    {
        DeviceGrid dv(v2.size());
        dv.moveToDevice(v2.v.data());
        v2 = Cube();  // release host memory

        RTMKernel kernel(p, plat);
        kernel.initKernel();
        if (p.modeling) {
            say(p, "Running RTM Modeling Process... ");
            int k = 0;
            for (const ShotPos &s : shots) {
                say(p, "Shot " + std::to_string(++k) + "/" + std::to_string(shots.size()) + " (" +
                           std::to_string(s.sx) + "," + std::to_string(s.sy) + "," +
                           std::to_string(s.sz) + "):");
                kernel.rtmModel(s, dv);
            }
        } else if (p.migration) {
            say(p, "Running RTM Migration Process... ");
            const int nxe = (int)p.nx + 2 * b, nye = (int)p.ny + 2 * b, nze = (int)p.nz + 2 * b;
            DeviceGrid total((size_t)nxe * nye * nze);
            int k = 0;
            for (const ShotPos &s : shots) {
                say(p, "Shot " + std::to_string(++k) + "/" + std::to_string(shots.size()) + " (" +
                           std::to_string(s.sx) + "," + std::to_string(s.sy) + "," +
                           std::to_string(s.sz) + "):");
                kernel.rtmMigrate(s, dv);
                hipcheck(rtmhip_stack(total.ptr(), kernel.img().ptr(), total.count()),
                         "image stack");
            }
            say(p, "Applying Laplacian Filter to output image... ");
            const std::vector<float> fc = stencil_coefs(2, p.dx, p.dy, p.dz);
            DeviceGrid out((size_t)p.nx * p.ny * p.nz);
            hipcheck(rtmhip_image_finish(total.ptr(), out.ptr(), nxe, nye, nze, b, fc.data()),
                     "image finish");
            std::vector<float> h(out.count());
            out.moveFromDevice(h.data());
            const std::string name =
                image_name(p.outdir, p.mname, (int)(p.source_count_x * p.source_count_y),
                           (int)p.nx, (int)p.ny, (int)p.nz, (int)p.nt, 1);
            say(p, "Saving rtm output image at: " + name);
            save_raw(name, h.data(), h.size());
        }
        plat.synchronize();
        if (p.verbose) std::cout << "> Time Report: " << std::endl
                                 << kernel.report().to_string() << std::endl;
    }
}

int run_json(const std::string &path, int device)
{
    try {
        RTMParam p;
        p.load(path);
        run(p, device);
    } catch (const RTMParamException &e) {
        std::cout << "> Error:\n" << e.what() << std::endl << "> Aborting!" << std::endl;
        return 1;
    } catch (const RTMException &e) {
        std::cout << "> Error:\n" << e.what() << std::endl << "> Aborting!" << std::endl;
        return 1;
    } catch (const std::exception &e) {
        std::cout << "> Error:\n" << e.what() << std::endl << "> Aborting!" << std::endl;
        return 1;
    }
    return 0;
}

}  // namespace rtm3d
