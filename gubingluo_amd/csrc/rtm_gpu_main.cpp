// rtm_gpu -- the drop-in driver executable (the reference's RTM_GPU_velstr.cpp
// main), linked against librtm_hip.so.
#include "rtm_hip.h"

int main(int argc, char** argv) { return rtmh_driver_main(argc, argv); }
