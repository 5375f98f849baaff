// platform.cpp -- RTMHIPPlatform and DeviceGrid over the C ABI.
#include "platform.hpp"

#include "param.hpp"

namespace rtm3d {

void hipcheck(int rc, const char *what)
{
    if (rc != RTMHIP_OK)
        throw RTMException(std::string(what) + ": " + rtmhip_last_error());
}

void DeviceGrid::createDeviceBuffer(size_t count)
{
    removeDeviceBuffer();
    void *p = nullptr;
    hipcheck(rtmhip_alloc(&p, count * sizeof(float)), "device allocation");
    p_ = static_cast<float *>(p);
    n_ = count;
    devMemSet(0.f);
}

void DeviceGrid::removeDeviceBuffer()
{
    if (p_) rtmhip_free(p_);
    p_ = nullptr;
    n_ = 0;
}

void DeviceGrid::moveToDevice(const float *host)
{
    hipcheck(rtmhip_copy_to_device(p_, host, n_ * sizeof(float)), "copy to device");
}

void DeviceGrid::moveFromDevice(float *host) const
{
    hipcheck(rtmhip_copy_from_device(host, p_, n_ * sizeof(float)), "copy from device");
}

void DeviceGrid::devMemSet(float value)
{
    hipcheck(rtmhip_fill(p_, n_, value), "device fill");
}

void RTMHIPPlatform::initRTMPlatform() { hipcheck(rtmhip_init(device_), "initRTMPlatform"); }

void RTMHIPPlatform::destroyRTMPlatform() { rtmhip_finalize(); }

void RTMHIPPlatform::rtmTaperAllBorders(float *P, const DeviceGrid &t, const Dims &d)
{
    hipcheck(rtmhip_taper(P, t.ptr(), d.nxe, d.nye, d.nze, d.blen, 0), "rtmTaperAllBorders");
}

void RTMHIPPlatform::rtmTaperUpperBorders(float *P, const DeviceGrid &t, const Dims &d)
{
    hipcheck(rtmhip_taper(P, t.ptr(), d.nxe, d.nye, d.nze, d.blen, 1), "rtmTaperUpperBorders");
}

void RTMHIPPlatform::rtmStep(const float *P, float *PP, const std::vector<float> &coefs,
                             const DeviceGrid &v2dt2, const Dims &d)
{
    hipcheck(rtmhip_step(P, PP, v2dt2.ptr(), coefs.data(), d.nxe, d.nye, d.nze, d.order),
             "rtmStep");
}

void RTMHIPPlatform::rtmImageCondition(float *img, const float *PS, const float *PR,
                                       const Dims &d)
{
    hipcheck(rtmhip_image_condition(img, PS, PR, d.nxe, d.nye, d.nze, d.blen),
             "rtmImageCondition");
}

void RTMHIPPlatform::rtmApplySource(float *PP, int sx, int sy, int sz, float value,
                                    const Dims &d)
{
    hipcheck(rtmhip_apply_source(PP, d.nxe, d.nye, d.nze, sx + d.blen, sy + d.blen, sz + d.blen,
                                 value),
             "rtmApplySource");
}

void RTMHIPPlatform::rtmSaveReceiverData(float *PP, float *rcv, const rtmhip_rcv_geom &g,
                                         int ntstep, int lt, const Dims &d)
{
    hipcheck(rtmhip_receivers(PP, rcv, d.nxe, d.nye, d.nze, d.blen, &g, ntstep, lt, 0),
             "rtmSaveReceiverData");
}

void RTMHIPPlatform::rtmRestoreReceiverData(float *PP, float *rcv, const rtmhip_rcv_geom &g,
                                            int ntstep, int lt, const Dims &d)
{
    hipcheck(rtmhip_receivers(PP, rcv, d.nxe, d.nye, d.nze, d.blen, &g, ntstep, lt, 1),
             "rtmRestoreReceiverData");
}

void RTMHIPPlatform::rtmSaveUpperBorder(float *PP, float *upb, int lt, const Dims &d)
{
    hipcheck(rtmhip_upper_border(PP, upb, d.nxe, d.nye, d.nze, d.blen, d.order, lt, 0),
             "rtmSaveUpperBorder");
}

void RTMHIPPlatform::rtmRestoreUpperBorder(float *PP, float *upb, int lt, const Dims &d)
{
    hipcheck(rtmhip_upper_border(PP, upb, d.nxe, d.nye, d.nze, d.blen, d.order, lt, 1),
             "rtmRestoreUpperBorder");
}

void RTMHIPPlatform::synchronize() { hipcheck(rtmhip_synchronize(), "synchronize"); }

}  // namespace rtm3d
