// Discrete adjoint of the 2D RHS and SSP-RK3 step (characteristic-wise
// WENO-JS + HLLC + MND6): device buffers and entry points.
#include <hip/hip_runtime.h>

#include "handle.hpp"

struct AdjointStore {
  int dummy;
};

void cfd_adjoint_release(cfd_handle* h) {
  if (!h->adj) return;
  delete h->adj;
  h->adj = nullptr;
}
