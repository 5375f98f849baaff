"""MI355X-native (gfx950) RTM propagation-and-imaging path.

Layout:
  csrc/            HIP kernels and C ABI (librtm3d_hip.so), C++ host driver
                   (librtm3d_host.so, bin/rtm3d)
  hip.py           ctypes binding of the C ABI (include/rtm_hip.h)
  host.py          ctypes binding of the host set-up library
  workloads.py     synthetic velocity models and shot geometries
  build.py         hipcc / g++ build of the above
"""
import os

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
REPO_DIR = os.path.dirname(PKG_DIR)
DRIVER = os.path.join(PKG_DIR, "bin", "rtm3d")

__version__ = "1.0.0"
