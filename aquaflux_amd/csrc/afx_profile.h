// In-library kernel timing: when enabled, every launch made through AFX_LAUNCH is bracketed by a
// pair of HIP events recorded on the launch stream, so per-kernel durations come from the stream
// the kernel actually ran on.  Disabled by default (a relaxed load of one flag per launch).
#pragma once

#include <hip/hip_runtime.h>

namespace afx {

enum KernelId : int {
  kKGeomFace = 0,
  kKGeomApprox,
  kKGeomOrient,
  kKGeomVolume,
  kKFlowGrad,
  kKFlowAp,
  kKFlowRes,
  kKFlowResAdj,
  kKFlowApAdj,
  kKFlowGradAdj,
  kKScalarGrad,
  kKScalarRes,
  kKScalarResAdj,
  kKScalarGradAdj,
  kKGeomAdj,
  kKSstClosure,
  kKSstRes,
  kKSstResAdj,
  kKSstClosureAdj,
  kKBcsrAssemble,
  kKBcsrSpmv,
  kKPbjSetup,
  kKPbjApply,
  kKBlas,
  kKRansGrad,
  kKRansAp,
  kKRansApAdj,
  kKRansGradAdj,
  kKScalarLimAdj,
  kKGeomAdjFace,
  kKGeomAdjNode,
  kKRansScalarRes,
  kKRansFlowResAdj,
  kKLimiter,
  kKCggSweep,
  kKCggSweepAdj,
  kKProbeScatter,
  kKAmg,
  kKKrylov,
  kKFaceFlux,
  kKNumKernels
};

extern const char* const kKernelNames[kKNumKernels];

// Returns an event pair to record around a launch, or false when profiling is off / pool full.
bool profile_begin(int id, hipStream_t s, hipEvent_t* stop);
void profile_end(hipEvent_t stop, hipStream_t s);

}  // namespace afx

#define AFX_LAUNCH(ID, KERNEL, GRID, BLOCK, SHMEM, STREAM, ...)                 \
  do {                                                                          \
    hipEvent_t _afx_stop = nullptr;                                             \
    bool _afx_prof = afx::profile_begin((ID), (STREAM), &_afx_stop);            \
    hipLaunchKernelGGL(KERNEL, GRID, BLOCK, SHMEM, STREAM, __VA_ARGS__);        \
    if (_afx_prof) afx::profile_end(_afx_stop, (STREAM));                       \
  } while (0)
