// access_check.h -- force-included (-include) into the host emulation when it
// is built for the sanitizer run: routes the kernels' field-access hook to a
// bounds check against the box allocations that the driver registered.
#pragma once
#ifdef __cplusplus
extern "C" {
#endif
void sspt_access_check(const void *base, unsigned long long byte_offset, unsigned long long bytes);
// buffers that the emulation allocates itself and the kernels index
void sspt_access_register(const void *base, unsigned long long bytes);
void sspt_access_unregister(const void *base);
#ifdef __cplusplus
}
#endif
#define SSPT_ACCESS_CHECK(p, off, n) sspt_access_check((p), (off), (n))
#define SSPT_ACCESS_REGISTER(p, n) sspt_access_register((p), (n))
#define SSPT_ACCESS_UNREGISTER(p) sspt_access_unregister((p))
