// Kernel launch layer. Every launch on the trajectory path goes through
// launch(): with no recorder attached it launches the single-chain kernel
// on the context stream; with a recorder attached (serving, see server.hip)
// it records the call instead. A batch of chains of one shape class records
// the same call sequence once per chain; the recorder then emits one launch
// per call of the batched kernel variant, whose argument table holds one
// entry per chain (blockIdx.z = chain). The emitted sequence has fixed
// arguments, so it can be captured in a HIP graph and replayed.
#pragma once
#include <cstring>
#include <string>
#include <vector>

#include "kernels.hpp"
#include "runtime.hpp"

namespace mag {

struct RecCall {
  const void* kb = nullptr;  // batched kernel entry, void(*)(const A*)
  dim3 grid;
  size_t argsz = 0;
  std::vector<char> blob;  // argsz bytes per recorded chain
  void* table = nullptr;   // device copy of blob (set when the plan is built)
};

struct Recorder {
  std::vector<RecCall> calls;
  size_t pos = 0;
  int chain = 0;
  bool mismatch = false;
  void begin_chain(int i) {
    chain = i;
    pos = 0;
  }
  void add(const void* kb, dim3 grid, const void* args, size_t argsz) {
    if (chain == 0) {
      RecCall c;
      c.kb = kb;
      c.grid = grid;
      c.argsz = argsz;
      calls.push_back(std::move(c));
    } else if (pos >= calls.size() || calls[pos].kb != kb || calls[pos].argsz != argsz ||
               calls[pos].grid.x != grid.x || calls[pos].grid.y != grid.y) {
      mismatch = true;
      return;
    }
    RecCall& c = calls[pos++];
    const size_t o = c.blob.size();
    c.blob.resize(o + argsz);
    std::memcpy(c.blob.data() + o, args, argsz);
  }
};

template <class A>
int launch(mag_ctx c, const char* timer, void (*ks)(A), void (*kb)(const A*), dim3 grid, const A& a) {
  if (c->rec) {
    c->rec->add((const void*)kb, grid, &a, sizeof(A));
    return MAG_OK;
  }
  KTimer t(c, timer);
  A copy = a;
  void* args[] = {&copy};
  HIPCHK(c, hipLaunchKernel((const void*)ks, grid, dim3(BLOCK), args, 0, c->stream));
  return MAG_OK;
}

// Device-to-device copy of `bytes` (a multiple of 16) through a kernel.
inline int copy_async(mag_ctx c, void* dst, const void* src, size_t bytes) {
  CopyArgs a{(double2*)dst, (const double2*)src, bytes / 16};
  const unsigned nb = (unsigned)std::min<size_t>((a.n + BLOCK - 1) / BLOCK, (size_t)1 << 30);
  return launch(c, "copy", k_copy<0>, k_copy_b<0>, dim3(nb ? nb : 1), a);
}

inline int fill_async(mag_ctx c, double* p, double v, int n) {
  FillArgs a{p, v, n};
  return launch(c, "fill", k_fill<0>, k_fill_b<0>, dim3((n + BLOCK - 1) / BLOCK), a);
}

}  // namespace mag
