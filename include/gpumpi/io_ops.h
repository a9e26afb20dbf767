// Requests of the device-to-host I/O channel (shared by the runtime and the device C
// library).  The reference left this delegation unimplemented
// (gpu_libs/gpu_libc/cuda_mpi.cu:474-510, gpu_main/device_host_comm.cu:5-15).
#pragma once

namespace gpumpi {
constexpr int kIoData = 4096;  // bytes each way per request
enum IoOp : int {
    kIoOpen = 1,   // data: path\0mode\0 -> handle (> 0) or 0
    kIoClose,      // a0 handle -> fclose
    kIoWrite,      // a0 handle, data[len] -> bytes written
    kIoRead,       // a0 handle, a1 max bytes -> bytes read into data
    kIoGets,       // a0 handle, a1 size -> length + 1 in data, -1 at end of file
    kIoGetc,       // a0 handle -> byte or -1
    kIoUngetc,     // a0 handle, a1 byte
    kIoEof,        // a0 handle
    kIoFlush,      // a0 handle
    kIoSeek,       // a0 handle, a1 offset, a2 whence
    kIoTell,       // a0 handle
};
// Handles of the standard streams (files opened by ranks get 1, 2, ...).
constexpr long long kIoStdin = -3, kIoStdout = -1, kIoStderr = -2;
}  // namespace gpumpi
