/* initialize.h -- device initialisation (util/cuda/initialize.h:4-10 of the reference).
 * ispm_initialize() aborts unless the current device is gfx950, the only ISA the library is
 * built for (the reference aborts on a compute-capability mismatch). */
#ifndef ISPM_INITIALIZE_H
#define ISPM_INITIALIZE_H

extern void ispm_initialize();

struct DeviceIDString {
  char str[32];
};

extern DeviceIDString get_device_string();

#endif
