// Host stand-in for <hip/hip_runtime.h> that runs device code with wave semantics on
// the CPU, for tests only.  The threads of a workgroup are fibers on one OS thread: a
// thread runs until its next cross-lane operation (__ballot, __shfl*, readlane,
// __syncthreads), then hands over to the next, so a cross-lane operation completes when
// the last live thread of the workgroup reaches it -- a round robin in which every
// thread must reach the same sequence of cross-lane operations, as a wave's lanes must
// on the GPU (ballots and shuffles read the thread's own wave of 64).  A thread that
// reaches a different operation than the others (divergent control flow around a
// cross-lane operation) stops the program with a message; threads that have returned
// take no part, like finished lanes.  Workgroups of one launch run on several OS
// threads at once (__shared__ variables are per OS thread), so device-scope atomics see
// real contention.  Each fiber stack has a guard page below it.
#pragma once
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <thread>
#include <vector>

#include <sys/mman.h>

#define __device__
#define __host__
#define __global__
#define __forceinline__ inline
#define __noinline__ __attribute__((noinline))
#define __launch_bounds__(x)
#define __shared__ static thread_local

using std::isfinite;

typedef int hipError_t;
typedef void* hipStream_t;
enum { hipSuccess = 0, hipErrorInvalidValue = 1, hipErrorNotSupported = 801 };
enum hipMemcpyKind { hipMemcpyHostToDevice, hipMemcpyDeviceToHost, hipMemcpyDeviceToDevice };
inline hipError_t hipGetLastError() { return hipSuccess; }
inline hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind, hipStream_t) {
  std::memcpy(d, s, n);
  return hipSuccess;
}
inline hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
typedef void* hipEvent_t;
inline hipError_t hipEventCreate(hipEvent_t*) { return hipSuccess; }
inline hipError_t hipEventDestroy(hipEvent_t) { return hipSuccess; }
inline hipError_t hipEventRecord(hipEvent_t, hipStream_t) { return hipSuccess; }
inline hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
inline hipError_t hipEventElapsedTime(float* ms, hipEvent_t, hipEvent_t) { *ms = 0.f; return hipSuccess; }
typedef void* hipGraph_t;
typedef void* hipGraphExec_t;
enum hipStreamCaptureMode { hipStreamCaptureModeGlobal, hipStreamCaptureModeThreadLocal };
inline hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode) { return hipErrorInvalidValue; }
inline hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t*) { return hipErrorInvalidValue; }
inline hipError_t hipGraphInstantiate(hipGraphExec_t*, hipGraph_t, void*, void*, size_t) { return hipErrorInvalidValue; }
inline hipError_t hipGraphLaunch(hipGraphExec_t, hipStream_t) { return hipErrorInvalidValue; }
inline hipError_t hipGraphExecDestroy(hipGraphExec_t) { return hipSuccess; }
inline hipError_t hipGraphDestroy(hipGraph_t) { return hipSuccess; }
// device queries: an emulated device of 4 compute units, one resident wave each
enum hipDeviceAttribute_t { hipDeviceAttributeMultiprocessorCount };
inline hipError_t hipGetDevice(int* d) { *d = 0; return hipSuccess; }
inline hipError_t hipDeviceGetAttribute(int* v, hipDeviceAttribute_t, int) { *v = 4; return hipSuccess; }
template <class F>
inline hipError_t hipOccupancyMaxActiveBlocksPerMultiprocessor(int* n, F, int, size_t) { *n = 1; return hipSuccess; }

struct dim3 {
  unsigned x, y, z;
  dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};
struct Idx3 { unsigned x = 0, y = 0, z = 0; };

template <class T> inline T max(T a, T b) { return a > b ? a : b; }
template <class T> inline T min(T a, T b) { return a < b ? a : b; }

// ---------------------------------------------------------------------------
// fibers: callee-saved registers, the FP control words and the stack pointer
// ---------------------------------------------------------------------------
extern "C" void simt_switch(void** save_sp, void* next_sp);
extern "C" void simt_entry();
asm(R"(
.text
.globl simt_switch
.type simt_switch,@function
simt_switch:
  pushq %rbp
  pushq %rbx
  pushq %r12
  pushq %r13
  pushq %r14
  pushq %r15
  subq $8, %rsp
  stmxcsr (%rsp)
  fnstcw 4(%rsp)
  movq %rsp, (%rdi)
  movq %rsi, %rsp
  ldmxcsr (%rsp)
  fldcw 4(%rsp)
  addq $8, %rsp
  popq %r15
  popq %r14
  popq %r13
  popq %r12
  popq %rbx
  popq %rbp
  ret
.size simt_switch,.-simt_switch
)");

namespace simt {
constexpr int WAVE = 64;
constexpr int MAXL = 1024;                 // threads of a workgroup
constexpr size_t STACK = 512u << 10;       // per fiber, plus a guard page
constexpr size_t PAGE = 4096;

struct Wave {
  int nlanes = 0;
  int cur = 0;
  void* lane_sp[MAXL];
  void* main_sp = nullptr;
  bool done[MAXL];
  int ndone = 0;
  unsigned long long seq[MAXL];      // cross-lane operations each thread has reached
  int site[MAXL];                    // which kind of operation (checked for divergence)
  uint64_t xch[2][MAXL];             // exchange buffers, alternating per operation
  const std::function<void()>* body = nullptr;
  char* stacks = nullptr;            // nstacks x (guard page + STACK), committed lazily
  int nstacks = 0;
  ~Wave() {
    if (stacks) munmap(stacks, (size_t)nstacks * (PAGE + STACK));
  }
};
inline thread_local Wave* W = nullptr;
inline thread_local Idx3 g_thread, g_block, g_bdim, g_gdim;

[[noreturn]] inline void die(const char* msg) {
  std::fprintf(stderr, "simt emulation: %s\n", msg);
  std::abort();
}

// the next lane that has not finished, after `from` (round robin)
inline int next_live(int from) {
  for (int k = 1; k <= W->nlanes; ++k) {
    const int l = (from + k) % W->nlanes;
    if (!W->done[l]) return l;
  }
  return -1;
}
inline void switch_to(int to) {
  const int me = W->cur;
  W->cur = to;
  g_thread.x = (unsigned)to;
  simt_switch(&W->lane_sp[me], W->lane_sp[to]);
  g_thread.x = (unsigned)W->cur;
}
// every live lane reaches this point before any passes it
inline void arrive(int site_id) {
  const int me = W->cur;
  W->site[me] = site_id;
  ++W->seq[me];
  const int nx = next_live(me);
  if (nx > me) {
    switch_to(nx);
  } else {
    // the last live lane in round-robin order: every live lane is here; check them
    for (int l = 0; l < W->nlanes; ++l)
      if (!W->done[l] && (W->seq[l] != W->seq[me] || W->site[l] != site_id))
        die("lanes of a wave reached different cross-lane operations (divergent control flow)");
    if (nx != me && nx >= 0) switch_to(nx);   // wake the first live lane
  }
}
// thread `src` took part in the cross-lane operation thread `me` has just completed
// (it reached that operation before returning, if it has returned)
inline bool took_part(int src, int me) { return src < W->nlanes && W->seq[src] >= W->seq[me]; }
inline uint64_t* exchange(uint64_t v, int site_id) {
  const int me = W->cur;
  const int buf = (int)(W->seq[me] & 1);
  W->xch[buf][me] = v;
  arrive(site_id);
  return W->xch[buf];
}
inline void lane_main() {
  (*W->body)();
  const int me = W->cur;
  // a live thread waiting at a cross-lane operation this thread never reached
  for (int l = 0; l < W->nlanes; ++l)
    if (!W->done[l] && l != me && W->seq[l] > W->seq[me])
      die("a thread returned while others wait at a cross-lane operation (divergent control flow)");
  W->done[me] = true;
  ++W->ndone;
  const int nx = next_live(me);
  if (nx < 0) {
    simt_switch(&W->lane_sp[me], W->main_sp);
  } else {
    W->cur = nx;
    g_thread.x = (unsigned)nx;
    simt_switch(&W->lane_sp[me], W->lane_sp[nx]);
  }
  die("a finished lane was resumed");
}
}  // namespace simt

// one translation unit per program includes this header (the asm symbols are global)
extern "C" __attribute__((used, noinline)) void simt_entry_c() { simt::lane_main(); }
asm(R"(
.text
.globl simt_entry
.type simt_entry,@function
simt_entry:
  andq $-16, %rsp
  call simt_entry_c
  ud2
.size simt_entry,.-simt_entry
)");

namespace simt {
// run one workgroup: nlanes fibers executing body, lane 0 first
inline void run_group(int nlanes, const std::function<void()>& body) {
  if (nlanes > MAXL) die("workgroup too large");
  thread_local Wave w;
  W = &w;
  w.nlanes = nlanes;
  w.body = &body;
  w.ndone = 0;
  if (nlanes > w.nstacks) {   // stacks for this many threads (kept for later groups)
    if (w.stacks) munmap(w.stacks, (size_t)w.nstacks * (PAGE + STACK));
    const size_t total = (size_t)nlanes * (PAGE + STACK);
    void* m = mmap(nullptr, total, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS | MAP_NORESERVE, -1, 0);
    if (m == MAP_FAILED) {
      std::perror("mmap");
      die("cannot map fiber stacks");
    }
    w.stacks = (char*)m;
    w.nstacks = nlanes;
    for (int l = 0; l < nlanes; ++l) mprotect(w.stacks + (size_t)l * (PAGE + STACK), PAGE, PROT_NONE);
  }
  for (int l = 0; l < nlanes; ++l) {
    w.done[l] = false;
    w.seq[l] = 0;
    w.site[l] = 0;
    // initial frame: simt_switch pops 8 bytes of control words, six registers, then
    // returns into simt_entry
    char* top = w.stacks + (size_t)(l + 1) * (PAGE + STACK);
    top = (char*)((uintptr_t)top & ~(uintptr_t)15);
    uint64_t* sp = (uint64_t*)top;
    *--sp = 0;                                   // alignment slot
    *--sp = (uint64_t)(uintptr_t)&simt_entry;    // return address
    for (int r = 0; r < 6; ++r) *--sp = 0;       // rbp rbx r12 r13 r14 r15
    --sp;
    uint32_t csr;
    uint16_t cw;
    asm volatile("stmxcsr %0" : "=m"(csr));
    asm volatile("fnstcw %0" : "=m"(cw));
    std::memcpy(sp, &csr, 4);
    std::memcpy((char*)sp + 4, &cw, 2);
    w.lane_sp[l] = sp;
  }
  w.cur = 0;
  g_thread.x = 0;
  simt_switch(&w.main_sp, w.lane_sp[0]);
  if (w.ndone != nlanes) die("a wave ended with lanes still running");
}
}  // namespace simt

#define threadIdx simt::g_thread
#define blockIdx simt::g_block
#define blockDim simt::g_bdim
#define gridDim simt::g_gdim

inline int __lane_id() { return simt::W->cur % simt::WAVE; }
inline void __syncthreads() { simt::arrive(1); }
inline unsigned long long __ballot(int p) {
  const int base = simt::W->cur / simt::WAVE * simt::WAVE;
  const uint64_t* x = simt::exchange(p ? 1u : 0u, 2);
  unsigned long long b = 0;
  for (int l = 0; l < simt::WAVE && base + l < simt::W->nlanes; ++l)
    if (simt::took_part(base + l, simt::W->cur) && x[base + l]) b |= 1ull << l;
  return b;
}
template <class T>
inline uint64_t simt_bits(T v) {
  static_assert(sizeof(T) <= 8, "one word");
  uint64_t u = 0;
  std::memcpy(&u, &v, sizeof(T));
  return u;
}
template <class T>
inline T simt_val(uint64_t u) {
  T v;
  std::memcpy(&v, &u, sizeof(T));
  return v;
}
template <class T> inline T __shfl(T v, int src, int width = 64) {
  const int me = simt::W->cur;
  const uint64_t* x = simt::exchange(simt_bits(v), 3);
  const int s = (me / width) * width + ((src % width) + width) % width;
  if (!simt::took_part(s, me)) simt::die("shuffle from a thread that has returned");
  return simt_val<T>(x[s]);
}
template <class T> inline T __shfl_xor(T v, int off, int width = 64) {
  const int me = simt::W->cur;
  const uint64_t* x = simt::exchange(simt_bits(v), 4);
  const int src = me ^ off;
  if (src / width != me / width || src >= simt::W->nlanes) return v;
  if (!simt::took_part(src, me)) simt::die("shuffle from a thread that has returned");
  return simt_val<T>(x[src]);
}
template <class T> inline T __shfl_down(T v, int off, int width = 64) {
  const int me = simt::W->cur;
  const uint64_t* x = simt::exchange(simt_bits(v), 5);
  const int src = me + off;
  return ((src / width) == (me / width) && simt::took_part(src, me)) ? simt_val<T>(x[src]) : v;
}
inline int simt_readlane(int v, int l) {
  const int base = simt::W->cur / simt::WAVE * simt::WAVE;
  const uint64_t* x = simt::exchange(simt_bits(v), 6);
  if (!simt::took_part(base + l, simt::W->cur)) simt::die("readlane of a thread that has returned");
  return simt_val<int>(x[base + l]);
}
#define __builtin_amdgcn_readlane(v, l) simt_readlane((v), (l))
// DPP source lanes for the controls the kernels use (quad_perm, row_mirror,
// row_half_mirror; every row and bank enabled, every lane of the wave taking part)
inline int simt_update_dpp(int old, int src, int ctrl, int row_mask, int bank_mask, bool bound_ctrl) {
  (void)old;
  (void)bound_ctrl;
  if (row_mask != 0xF || bank_mask != 0xF) simt::die("dpp masks not emulated");
  const int me = simt::W->cur, lane = me % simt::WAVE, base = me - lane;
  const uint64_t* x = simt::exchange(simt_bits(src), 7);
  const int row = lane & ~15, i = lane & 15;
  int s;
  if (ctrl < 0x100) s = (lane & ~3) + ((ctrl >> (2 * (lane & 3))) & 3);
  else if (ctrl == 0x140) s = row + 15 - i;
  else if (ctrl == 0x141) s = row + (i & 8) + 7 - (i & 7);
  else simt::die("dpp control not emulated");
  if (base + s >= simt::W->nlanes || !simt::took_part(base + s, me)) simt::die("dpp from a lane that has returned");
  return simt_val<int>(x[base + s]);
}
#define __builtin_amdgcn_update_dpp(old, src, ctrl, rm, bm, bc) \
  simt_update_dpp((old), (src), (ctrl), (rm), (bm), (bc))
// the learned predictor's matrix-core products are not emulated (its kernels are not
// launched here)
#define __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, x, y, z) (c)
// v_mfma_{f64,f32}_16x16x4: D = C + A B over one 16 x 16 x 4 block, with the operand
// layout measured on gfx950: lane l holds A[l % 16][l / 16] and B[l / 16][l % 16]; its
// register v holds D[l / 16 + 4 v][l % 16] (FP64) or D[4 (l / 16) + v][l % 16] (FP32).
// The products are added to C in k order.  Each lane copies the exchanged operands
// before its next cross-lane operation.
inline long simt_mfma_calls = 0;   // products run (per lane), for the tests
template <class T, class V4>
inline V4 simt_mfma_16x16x4(T a, T b, V4 c) {
  ++simt_mfma_calls;
  const int me = simt::W->cur;
  const int base = me / simt::WAVE * simt::WAVE;
  T A[64], B[64];
  const uint64_t* xa = simt::exchange(simt_bits(a), 9);
  for (int l = 0; l < 64; ++l) {
    if (!simt::took_part(base + l, me)) simt::die("a matrix-core operation needs all 64 lanes");
    A[l] = simt_val<T>(xa[base + l]);
  }
  const uint64_t* xb = simt::exchange(simt_bits(b), 10);
  for (int l = 0; l < 64; ++l) B[l] = simt_val<T>(xb[base + l]);
  const int lane = me - base, j = lane % 16, ig = lane / 16;
  V4 d = c;
  for (int v = 0; v < 4; ++v) {
    const int i = sizeof(T) == 8 ? ig + 4 * v : 4 * ig + v;
    T acc = c[v];
    for (int k = 0; k < 4; ++k) acc += A[i + 16 * k] * B[j + 16 * k];
    d[v] = acc;
  }
  return d;
}
#define __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, x, y, z) simt_mfma_16x16x4<double>((a), (b), (c))
#define __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, x, y, z) simt_mfma_16x16x4<float>((a), (b), (c))

inline int __ffsll(long long x) { return __builtin_ffsll(x); }
inline long long __double_as_longlong(double x) { long long r; std::memcpy(&r, &x, 8); return r; }
inline double __longlong_as_double(long long x) { double r; std::memcpy(&r, &x, 8); return r; }
inline int __popcll(unsigned long long x) { return __builtin_popcountll(x); }
inline unsigned int __float_as_uint(float x) {
  unsigned int u;
  __builtin_memcpy(&u, &x, 4);
  return u;
}
template <class T> inline T atomicAdd(T* p, T v) { return __atomic_fetch_add(p, v, __ATOMIC_RELAXED); }
template <class T> inline T atomicMax(T* p, T v) {
  T o = __atomic_load_n(p, __ATOMIC_RELAXED);
  while (v > o && !__atomic_compare_exchange_n(p, &o, v, false, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {
  }
  return o;
}

// Workgroups of a launch are spread over up to SIMT_THREADS OS threads; within one,
// lanes are fibers.
#ifndef SIMT_THREADS
#define SIMT_THREADS 8
#endif
#define hipLaunchKernelGGL(kernel, grid, block, shmem, stream, ...)                           \
  do {                                                                                       \
    const dim3 g_ = (grid), b_ = (block);                                                    \
    std::atomic<unsigned> next_{0};                                                          \
    auto worker_ = [&]() {                                                                   \
      for (unsigned bx_ = next_++; bx_ < g_.x; bx_ = next_++) {                              \
        simt::g_bdim.x = b_.x;                                                               \
        simt::g_gdim.x = g_.x;                                                               \
        simt::g_block.x = bx_;                                                               \
        const std::function<void()> body_ = [&]() { kernel(__VA_ARGS__); };                  \
        simt::run_group((int)b_.x, body_);                                                   \
      }                                                                                      \
    };                                                                                       \
    const unsigned nt_ = g_.x < (unsigned)SIMT_THREADS ? g_.x : (unsigned)SIMT_THREADS;      \
    std::vector<std::thread> th_;                                                            \
    for (unsigned t_ = 1; t_ < nt_; ++t_) th_.emplace_back(worker_);                         \
    worker_();                                                                               \
    for (auto& t_ : th_) t_.join();                                                          \
  } while (0)
