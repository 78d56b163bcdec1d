// Helpers of the reference-side binding (fip_hip_seam.cpp) that the edited src/main.cpp
// calls in place of its cudaMemcpy / cudaMemcpy2D / cudaDeviceReset calls.
#pragma once

// One value of the current ("1") buffers at cell (x, y): field FIP_FIELD_H, _QX or _QY
// (src/main.cpp:53-83, sample()).
float fipHipCell(int field, int x, int y);
// A whole W x H field into dst (unpadded, x fastest): dumpDeviceBuffers() takes the "2"
// buffers (FIP_SIDE_OTHER) of h, qx, qy and hmax (src/main.cpp:180-202).
void fipHipDownload(int field, int side, float* dst);
// Releases the simulation (in place of cudaDeviceReset(), src/main.cpp:256).
void fipHipShutdown();
