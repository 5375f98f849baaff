// main.cpp -- rtm3d command line: `rtm3d input.json` (rtm-3d/src/Main.cpp:11-48).
// The optional second argument of the reference (an FPGA xclbin) is accepted
// and ignored.  The device is taken from RTM3D_DEVICE (default 0).
#include <cstdlib>
#include <iostream>

#include "controller.hpp"

int main(int argc, char **argv)
{
    if (argc < 2) {
        std::cout << "> You forgot the input.json file!" << std::endl;
        return -1;
    }
    const char *dev = std::getenv("RTM3D_DEVICE");
    return rtm3d::run_json(argv[1], dev ? std::atoi(dev) : 0);
}
