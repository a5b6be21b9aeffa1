// Instantiates the quartet kernels whose bra pair class is (dp|; one translation unit per bra class so they compile in parallel.
#include "vsvb_eri_kernels.h"

namespace vsvb {

hipError_t launch_quartet_b4(int kc, const QuartetArgs& a, bool jk, bool tr, int nblocks, size_t lds, hipStream_t st) {
  switch (kc) {
    case 0: return launch_q<2, 1, 0, 0>(a, jk, tr, nblocks, lds, st);
    case 1: return launch_q<2, 1, 1, 0>(a, jk, tr, nblocks, lds, st);
    case 2: return launch_q<2, 1, 1, 1>(a, jk, tr, nblocks, lds, st);
    case 3: return launch_q<2, 1, 2, 0>(a, jk, tr, nblocks, lds, st);
    case 4: return launch_q<2, 1, 2, 1>(a, jk, tr, nblocks, lds, st);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_groups_b4(int kc, const QuartetArgs* gs, int ng, bool tr, int nblocks, size_t lds,
                            unsigned long long* stamp, hipStream_t st) {
  switch (kc) {
    case 0: return launch_q_groups<2, 1, 0, 0>(gs, ng, tr, nblocks, lds, stamp, st);
    case 1: return launch_q_groups<2, 1, 1, 0>(gs, ng, tr, nblocks, lds, stamp, st);
    case 2: return launch_q_groups<2, 1, 1, 1>(gs, ng, tr, nblocks, lds, stamp, st);
    case 3: return launch_q_groups<2, 1, 2, 0>(gs, ng, tr, nblocks, lds, stamp, st);
    case 4: return launch_q_groups<2, 1, 2, 1>(gs, ng, tr, nblocks, lds, stamp, st);
    default: return hipErrorInvalidValue;
  }
}

int occupancy_b4(int kc, bool jk, bool tr, size_t lds) {
  switch (kc) {
    case 0: return occupancy_q<2, 1, 0, 0>(jk, tr, lds);
    case 1: return occupancy_q<2, 1, 1, 0>(jk, tr, lds);
    case 2: return occupancy_q<2, 1, 1, 1>(jk, tr, lds);
    case 3: return occupancy_q<2, 1, 2, 0>(jk, tr, lds);
    case 4: return occupancy_q<2, 1, 2, 1>(jk, tr, lds);
    default: return 1;
  }
}

size_t scratch_b4(int kc, bool jk, bool tr) {
  switch (kc) {
    case 0: return scratch_q<2, 1, 0, 0>(jk, tr);
    case 1: return scratch_q<2, 1, 1, 0>(jk, tr);
    case 2: return scratch_q<2, 1, 1, 1>(jk, tr);
    case 3: return scratch_q<2, 1, 2, 0>(jk, tr);
    case 4: return scratch_q<2, 1, 2, 1>(jk, tr);
    default: return 0;
  }
}

hipError_t launch_schwarz_b4(Pair* pairs, const PrimPair* pp, const double* boys, int p0, int np, hipStream_t st) {
  if (np <= 0) return hipSuccess;
  hipLaunchKernelGGL((schwarz_kernel<2, 1>), dim3((np + 3) / 4), dim3(256), 0, st, pairs, pp, boys, p0, np);
  return hipGetLastError();
}

}  // namespace vsvb
