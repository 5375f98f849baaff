// platform.hpp -- the reference's operator interface on top of the C ABI.
//
// RTMHIPPlatform provides the operators of class RTMPlatform
// (rtm-3d/include/RTMPlatform.hpp:22-114) with the same names and argument
// meaning; DeviceGrid is the device-buffer slot of the reference's grids
// (rtm-3d/include/RTMDeviceGrid.hpp:19-62).  A failing call throws
// RTMException carrying the library's message, which the driver turns into
// exit status 1 as the reference does.
#pragma once
#include <cstddef>
#include <string>
#include <vector>

#include "rtm_hip.h"

namespace rtm3d {

void hipcheck(int rc, const char *what);

// Owns one device buffer of `count` floats.
class DeviceGrid {
public:
    DeviceGrid() = default;
    explicit DeviceGrid(size_t count) { createDeviceBuffer(count); }
    ~DeviceGrid() { removeDeviceBuffer(); }
    DeviceGrid(const DeviceGrid &) = delete;
    DeviceGrid &operator=(const DeviceGrid &) = delete;
    DeviceGrid(DeviceGrid &&o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr, o.n_ = 0; }

    void createDeviceBuffer(size_t count);
    void removeDeviceBuffer();
    void moveToDevice(const float *host);
    void moveFromDevice(float *host) const;
    void devMemSet(float value);
    float *ptr() const { return p_; }
    size_t count() const { return n_; }

private:
    float *p_ = nullptr;
    size_t n_ = 0;
};

struct Dims {
    int nxe, nye, nze, blen, order;
};

class RTMHIPPlatform {
public:
    explicit RTMHIPPlatform(int device = 0) : device_(device) {}
    void initRTMPlatform();
    void destroyRTMPlatform();

    void rtmTaperAllBorders(float *P, const DeviceGrid &taper, const Dims &d);
    void rtmTaperUpperBorders(float *P, const DeviceGrid &taper, const Dims &d);
    void rtmStep(const float *P, float *PP, const std::vector<float> &coefs,
                 const DeviceGrid &v2dt2, const Dims &d);
    void rtmImageCondition(float *img, const float *PS, const float *PR, const Dims &d);
    void rtmApplySource(float *PP, int sx, int sy, int sz, float value, const Dims &d);
    void rtmSaveReceiverData(float *PP, float *rcv, const rtmhip_rcv_geom &g, int ntstep, int lt,
                             const Dims &d);
    void rtmRestoreReceiverData(float *PP, float *rcv, const rtmhip_rcv_geom &g, int ntstep,
                                int lt, const Dims &d);
    void rtmSaveUpperBorder(float *PP, float *upb, int lt, const Dims &d);
    void rtmRestoreUpperBorder(float *PP, float *upb, int lt, const Dims &d);
    void synchronize();

private:
    int device_;
};

}  // namespace rtm3d
