/* include/gpu/gpu.h:8 of the reference: 1 when a gfx950 device is usable. */
#ifndef HPRQP_GPU_H
#define HPRQP_GPU_H

#ifdef __cplusplus
extern "C" {
#endif

int gpu_backend_available(void);

#ifdef __cplusplus
}
#endif

#endif
