// Optional per-kernel timing with HIP events, recorded on the launch stream around every
// kernel the library enqueues while profiling is on (gdx_profile_begin / _end).
#pragma once
#include <hip/hip_runtime.h>

namespace gdx {

bool prof_active();
// Records a start event; returns an opaque slot (or -1 when profiling is off).
int prof_start(const char* name, hipStream_t st);
void prof_stop(int slot, hipStream_t st);

struct ProfMark {
  int slot;
  hipStream_t st;
  ProfMark(const char* name, hipStream_t s) : slot(prof_active() ? prof_start(name, s) : -1), st(s) {}
  void stop() {
    if (slot >= 0) prof_stop(slot, st);
  }
};

}  // namespace gdx

#define GDX_LAUNCH(name, kern, grid, block, shm, st, ...)                 \
  do {                                                                    \
    ::gdx::ProfMark _pm(name, st);                                        \
    hipLaunchKernelGGL(kern, grid, block, shm, st, __VA_ARGS__);          \
    _pm.stop();                                                           \
  } while (0)
