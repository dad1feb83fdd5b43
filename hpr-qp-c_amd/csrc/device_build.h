// Model preprocessing of build_from_QAbc on the GPU (src/shared/build.c:157-337,
// src/shared/sparse.c:81-125): the symmetrisation of Q and the transpose of A
// for large models. Both return exactly the host results; NOT_IMPLEMENTED means
// the device path does not apply (no gfx950 device, a small model, indices
// beyond int32 or out of range, HPRQP_HOST_BUILD set) and the caller runs the
// host path instead.
#pragma once

#include <cstddef>

#include "shared/types.h"

namespace hprqp {

// Q_sym = (Q + Q') / 2 with the reference's entry order and drop rule into a
// freshly allocated host CSR `out`.
HPRQPStatus device_symmetrize_q(const HPRQPMatrixCSR *q, HPRQPMatrixCSR *out);
// at = a' (stable: rows ascending within each column) into a freshly
// allocated host CSR.
HPRQPStatus device_transpose_csr(const HPRQPMatrixCSR *a, HPRQPMatrixCSR *at);

}  // namespace hprqp
