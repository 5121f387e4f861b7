// rhg_internal.h -- library-internal entry points used by the screening
// service (rhg_serve.cpp); not part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/rhg.h"

// Device memory shared by many contexts (slabs carved and reused; freed
// only by rhg_pool_destroy, after every context that used it is destroyed).
struct rhg_pool;
rhg_pool *rhg_pool_create(int device, size_t slab_bytes);
void rhg_pool_destroy(rhg_pool *p);

// rhg_create on a stream the caller owns (the context never destroys it);
// with ws non-null the context is attached to that workspace from the start,
// so that its set-up scratch lives there; with pool non-null its tables and
// own buffers come from the pool (and go back to it when it is destroyed).
int rhg_create_on_stream(const rhg_framework *fw, const rhg_cell *cell, const rhg_forcefield *ff, int device,
                         hipStream_t stream, rhg_workspace *ws, rhg_pool *pool, rhg_ctx **out);
// Moves a context to another stream once its current stream is idle;
// stream = null leaves it without one (no call may then be made on it but
// rhg_use_stream and rhg_destroy), so that destroying it waits for nothing.
int rhg_use_stream(rhg_ctx *c, hipStream_t stream);
