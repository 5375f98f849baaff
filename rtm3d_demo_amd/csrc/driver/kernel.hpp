// kernel.hpp -- per-shot time loops of the RTM path on the device.
//
// Mirrors RTMFiniteDifferencesKernel::rtmModel (rtm-3d/src/kernel/
// RTMFiniteDifferencesKernel.cpp:103-246) and RTMHBCKernel / RTMRBCKernel
// ::rtmMigrate (rtm-3d/src/kernel/RTMFiniteDifferencesKernel_HBC.cpp:64-365,
// ..._RBC.cpp:36-274) and RTMFreqDomainKernel::rtmMigrate
// (rtm-3d/src/kernel/RTMFiniteDifferencesKernel_FreqDomain.cpp:207-369):
// same operator order, same chunked seismogram files,
// same outputs.  Differences by design: wavefields, the upper-border history
// and the receiver chunks stay in device memory for the whole shot (no per
// operator synchronisation, no UPB spill files), and the receiver step of the
// backward pass is fused with receiver injection and imaging.
#pragma once
#include <string>
#include <vector>

#include "grid.hpp"
#include "param.hpp"
#include "platform.hpp"

namespace rtm3d {

struct ShotPos {
    int sx, sy, sz;  // interior coordinates
};

struct KernelReport {
    double modeling_s = 0, forward_s = 0, backward_s = 0, migration_s = 0;
    long long modeling_n = 0, migration_n = 0;
    double cell_updates = 0;  // wavefield cell updates performed
    std::string to_string() const;
};

class RTMKernel {
public:
    RTMKernel(const RTMParam &p, RTMHIPPlatform &plat);
    ~RTMKernel();
    void initKernel();
    void rtmModel(const ShotPos &s, const DeviceGrid &v2dt2);
    // Migrates one shot; the shot image (extended grid) is left in img().
    void rtmMigrate(const ShotPos &s, const DeviceGrid &v2dt2);
    const DeviceGrid &img() const { return img_; }
    KernelReport &report() { return report_; }

private:
    const RTMParam &p_;
    RTMHIPPlatform &plat_;
    Dims d_{};
    bool hbc_ = true;
    int wave_format_ = RTMHIP_WAVE_F32;  // RTM3D_WAVE_FORMAT
    int backward_mode_ = 1;              // RTM3D_BACKWARD_MODE
    // HBC upper-border history: every step on the device, or (spill) one
    // receiver chunk on the device and the whole history in host memory
    bool upb_spill_ = false;             // RTM3D_UPB_SPILL, or no room on the device
    size_t upb_chunk_bytes_ = 0;         // one chunk's layers
    std::vector<unsigned char> host_upb_;
    std::vector<float> coefs_;
    DeviceGrid taper_, ricker_, rcv_, upb_, img_;
    DeviceGrid f_[4];  // two source and two receiver wavefields
    KernelReport report_;
    std::vector<float> host_rcv_;
    // frequency-domain migration (fmig)
    FreqTables freq_;
    int nwstep_ = 0;
    DeviceGrid kre_, kim_, w2_, acc_[4];  // accumulators sRe, sIm, rRe, rIm
    void rtmMigrateFreq(const ShotPos &s, const DeviceGrid &v2dt2);

    void read_options();
    rtmhip_shot make_shot(const ShotPos &s) const;
    rtmhip_rcv_geom geom() const;
    void snapshot(const float *field, int sx, const std::string &file);
    void progress(const char *tag, const ShotPos &s, int it, double t0);
};

}  // namespace rtm3d
