// Uneven batched MFT: many (field, wavelength, shift) instances of different
// pupil and output sizes, computed by grouped persistent launches (gfx950).
//
// One call of dlx_mft_batched_psf runs, whatever the mix of shapes:
//   1. one launch that converts every distinct complex64 input field to split4;
//   2. one launch that generates every instance's transfer matrices W_x, W_y
//      (split4, FP64 phase arguments);
//   3. one grouped GEMM launch for all stage-1 products U_i^T = (P_i W_x,i^T)^T;
//   4. one grouped GEMM launch for all stage-2 products with the |E|^2
//      epilogue, I_i = |alpha_i W_y,i U_i|^2 (alpha_i = c_i sqrt(weight_i));
//   5. one launch that sums the planes of each output image, in a fixed order.
// Per-job descriptors live in the caller's workspace; the host fills them in
// a pinned staging slot and copies them with the launches (stream-ordered).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <cstdlib>
#include <map>
#include <mutex>
#include <numeric>
#include <vector>

#include "dlx_internal.h"
#include "dlx_split.h"

namespace dlx {
namespace {

// ---------------------------------------------------------------------------
// device job descriptors
// ---------------------------------------------------------------------------
struct CvtJob {              // complex64 rows x k -> split4
  const float2* in;
  uint8_t* out;
  long row_start;            // first row of this job in the launch's row list
  int rows, k;
};

struct MatJob {              // split4 rows x cols of exp(sign 2 pi i u_r v_c)
  uint8_t* out;
  long row_start;
  int rows, cols, sign_inv, pad_;
  double c_scale, c_off;     // column coordinate v_c = c * c_scale + c_off
  double r_scale, r_off;     // row coordinate    u_r = r * r_scale + r_off
};

struct SumJob {              // out[p] = sum_{q < count} planes[q * npix + p]
  float* out;
  const float* planes;
  long npix;
  int count, pad_;
};

// Workgroup-per-row kernels: a workgroup takes whole rows (grid-stride over the
// launch's row list), finds the row's job once with a workgroup-uniform binary
// search on scalar loads, and its lanes cover the row's chunks, a quad of
// lanes per chunk with quad-cooperative 16-byte stores.  Chunk counts times 4
// and the workgroup size are multiples of 4, so whole quads take each chunk.
template <typename Job>
__device__ __forceinline__ int find_row_job(const Job* jobs, int njobs, long row) {
  int lo = 0, hi = njobs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (jobs[mid].row_start <= row) lo = mid; else hi = mid - 1;
  }
  return lo;
}

__global__ void __launch_bounds__(256)
group_to_split4_kernel(const CvtJob* __restrict__ jobs, int njobs, long total_rows) {
  for (long row = blockIdx.x; row < total_rows; row += gridDim.x) {
    const CvtJob& j = jobs[find_row_job(jobs, njobs, row)];
    const long r = row - j.row_start;
    const int kc = (int)split4_chunks(j.k);
    const float2* src = j.in + r * j.k;
    for (int item = threadIdx.x; item < 4 * kc; item += blockDim.x) {
      const int q = item & 3, ch = item >> 2;
      const int x = ch * 8 + 2 * q;
      const float2 v0 = x < j.k ? src[x] : make_float2(0.f, 0.f);
      const float2 v1 = x + 1 < j.k ? src[x + 1] : make_float2(0.f, 0.f);
      store_chunk_quad(j.out, r, split4_row_bytes(j.k), ch, q, v0, v1);
    }
  }
}

__global__ void __launch_bounds__(256)
group_dft_split4_kernel(const MatJob* __restrict__ jobs, int njobs, long total_rows) {
  for (long row = blockIdx.x; row < total_rows; row += gridDim.x) {
    const MatJob& j = jobs[find_row_job(jobs, njobs, row)];
    const long r = row - j.row_start;
    const int kc = (int)split4_chunks(j.cols);
    const double u = (double)r * j.r_scale + j.r_off;
    for (int item = threadIdx.x; item < 4 * kc; item += blockDim.x) {
      const int q = item & 3, ch = item >> 2;
      float2 v[2];
#pragma unroll
      for (int e = 0; e < 2; ++e) {
        const int c = ch * 8 + 2 * q + e;
        if (c < j.cols) {
          double t = u * ((double)c * j.c_scale + j.c_off);
          t -= rint(t);
          // hardware sin/cos take turns; max abs error 1.3e-7 on [-1/2, 1/2]
          const float sn = __builtin_amdgcn_sinf((float)t), cs = __builtin_amdgcn_cosf((float)t);
          v[e] = make_float2(cs, j.sign_inv ? sn : -sn);
        } else {
          v[e] = make_float2(0.f, 0.f);
        }
      }
      store_chunk_quad(j.out, r, split4_row_bytes(j.cols), ch, q, v[0], v[1]);
    }
  }
}

__global__ void __launch_bounds__(256)
group_sum_kernel(const SumJob* __restrict__ jobs) {
  const SumJob& j = jobs[blockIdx.y];
  for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < j.npix;
       p += (long)gridDim.x * blockDim.x) {
    float acc = 0.f;
    for (int q = 0; q < j.count; ++q) acc += j.planes[(long)q * j.npix + p];
    j.out[p] = acc;
  }
}

// ---------------------------------------------------------------------------
// pinned staging slots for descriptor uploads
// ---------------------------------------------------------------------------
struct Slot {
  void* host = nullptr;
  size_t cap = 0;
  hipEvent_t done = nullptr;
  bool used = false;
};

struct Staging {
  std::mutex mu;
  static constexpr int kSlots = 16;
  Slot slots[kSlots];
  int next = 0;
};

Staging& staging() {
  static Staging s;
  return s;
}

// Copy `bytes` of host descriptors to device `dst` on stream s through a
// pinned slot; the slot is reused only after the copy that last used it ran.
int upload(const void* src, size_t bytes, void* dst, hipStream_t s) {
  if (bytes == 0) return DLX_OK;
  Staging& st = staging();
  std::lock_guard<std::mutex> lock(st.mu);
  Slot& sl = st.slots[st.next];
  st.next = (st.next + 1) % Staging::kSlots;
  if (sl.used) DLX_HIP(hipEventSynchronize(sl.done));
  if (sl.cap < bytes) {
    if (sl.host) DLX_HIP(hipHostFree(sl.host));
    sl.host = nullptr;
    sl.cap = 0;
    size_t cap = 1 << 16;
    while (cap < bytes) cap <<= 1;
    DLX_HIP(hipHostMalloc(&sl.host, cap, hipHostMallocDefault));
    sl.cap = cap;
  }
  if (!sl.done) DLX_HIP(hipEventCreateWithFlags(&sl.done, hipEventDisableTiming));
  std::memcpy(sl.host, src, bytes);
  DLX_HIP(hipMemcpyAsync(dst, sl.host, bytes, hipMemcpyHostToDevice, s));
  DLX_HIP(hipEventRecord(sl.done, s));
  sl.used = true;
  return DLX_OK;
}

// ---------------------------------------------------------------------------
// planning
// ---------------------------------------------------------------------------
struct Plan {
  // workspace offsets (bytes)
  size_t off_cvt, off_mat, off_g1, off_g2, off_sum, off_P, off_W, off_U, off_I, total;
  std::vector<CvtJob> cvt;
  std::vector<MatJob> mat;
  std::vector<GroupProblem> g1, g2, g1n, g2n;   // n: problems with N <= 64 (128 x 64 tiles)
  std::vector<SumJob> sum;
  int g1_xcd[9], g2_xcd[9], g1n_xcd[9], g2n_xcd[9];
  long cvt_work = 0, mat_work = 0;   // rows of the conversion / matrix launches
  std::vector<int> field_of;       // instance -> distinct-field index
  std::vector<const void*> fields;
  std::vector<int> field_n;
  std::vector<size_t> P_off, Wx_off, Wy_off, U_off, I_off;
};

int validate(int n, const dlx_mft_instance* inst, int n_outputs) {
  DLX_CHECK_ARG(n >= 0, "n must be >= 0");
  DLX_CHECK_ARG(n == 0 || inst, "instances is NULL");
  DLX_CHECK_ARG(n_outputs >= 0 && n_outputs <= 65535, "n_outputs must be in [0, 65535]");
  for (int i = 0; i < n; ++i) {
    const dlx_mft_instance& d = inst[i];
    DLX_CHECK_ARG(d.n_in > 0 && d.n_in <= 16384 && d.n_out > 0 && d.n_out <= 16384,
                  "instance %d: n_in and n_out must be in [1, 16384]", i);
    DLX_CHECK_ARG(d.field != nullptr, "instance %d: field is NULL", i);
    DLX_CHECK_ARG(d.out_index >= 0 && d.out_index < n_outputs,
                  "instance %d: out_index %d outside [0, %d)", i, d.out_index, n_outputs);
    DLX_CHECK_ARG(std::isfinite(d.wavelength) && d.wavelength > 0,
                  "instance %d: wavelength must be > 0", i);
    DLX_CHECK_ARG(std::isfinite(d.pixel_scale_in) && d.pixel_scale_in != 0,
                  "instance %d: pixel_scale_in must be non-zero", i);
    DLX_CHECK_ARG(std::isfinite(d.pixel_scale_out), "instance %d: pixel_scale_out", i);
    DLX_CHECK_ARG(std::isfinite(d.shift_x) && std::isfinite(d.shift_y),
                  "instance %d: shifts must be finite", i);
    DLX_CHECK_ARG(std::isfinite(d.weight) && d.weight >= 0, "instance %d: weight must be >= 0",
                  i);
    if (d.has_focal)
      DLX_CHECK_ARG(std::isfinite(d.focal_length + d.focal_shift) &&
                        d.focal_length + d.focal_shift != 0.0,
                    "instance %d: focal_length + focal_shift must be finite and non-zero", i);
  }
  std::vector<int> nout(n_outputs, 0);
  for (int i = 0; i < n; ++i) {
    int& o = nout[inst[i].out_index];
    DLX_CHECK_ARG(o == 0 || o == inst[i].n_out,
                  "instances of output %d have different n_out (%d and %d)",
                  inst[i].out_index, o, inst[i].n_out);
    o = inst[i].n_out;
  }
  return DLX_OK;
}

dlx_mft_geom geom_of(const dlx_mft_instance& d) {
  dlx_mft_geom g{};
  g.n_in = d.n_in;
  g.n_out = d.n_out;
  g.inverse = d.inverse;
  g.has_focal = d.has_focal;
  g.focal_length = d.focal_length;
  g.focal_shift = d.focal_shift;
  return g;
}

// Deal problems (largest first) round-robin to the 8 XCD ranges, then lay the
// ranges out contiguously and number their tiles.
void order_problems(std::vector<GroupProblem>& probs, int (&xcd)[9]) {
  std::vector<size_t> idx(probs.size());
  std::iota(idx.begin(), idx.end(), 0);
  auto cost = [&](const GroupProblem& p) {
    return (double)p.tiles_m * p.tiles_n * (double)split4_chunks(p.K);
  };
  std::stable_sort(idx.begin(), idx.end(),
                   [&](size_t a, size_t b) { return cost(probs[a]) > cost(probs[b]); });
  std::vector<std::vector<GroupProblem>> lists(8);
  std::vector<double> load(8, 0.0);
  for (size_t q : idx) {
    // greedy: the XCD with the least work so far
    int best = 0;
    for (int x = 1; x < 8; ++x)
      if (load[x] < load[best]) best = x;
    lists[best].push_back(probs[q]);
    load[best] += cost(probs[q]);
  }
  probs.clear();
  int t = 0;
  for (int x = 0; x < 8; ++x) {
    xcd[x] = t;
    for (GroupProblem& p : lists[x]) {
      p.tile_start = t;
      t += p.tiles_m * p.tiles_n;
      probs.push_back(p);
    }
  }
  xcd[8] = t;
}

GroupProblem make_problem(const uint8_t* A, long arb, const uint8_t* B, long brb, void* C,
                          long ldc, int M, int N, int K, float alpha) {
  GroupProblem p{};
  p.A = A; p.a_row_bytes = arb;
  p.B = B; p.b_row_bytes = brb;
  p.C = C; p.ldc = ldc;
  p.M = M; p.N = N; p.K = K;
  const int tile_n = N <= 64 ? 64 : 128;
  p.tiles_m = (M + 127) / 128;
  p.tiles_n = (N + tile_n - 1) / tile_n;
  p.alpha = alpha;
  return p;
}

// Problems with N <= 64 go to a launch with 128 x 64 tiles (no half-empty tiles).
void split_narrow(std::vector<GroupProblem>& all, std::vector<GroupProblem>& narrow) {
  std::vector<GroupProblem> wide;
  for (const GroupProblem& p : all) (p.N <= 64 ? narrow : wide).push_back(p);
  all.swap(wide);
}

// Sizes and offsets; with ws != nullptr also the device pointers in the jobs.
int plan_batched(int n, const dlx_mft_instance* inst, int n_outputs, float* const* outs,
                 uint8_t* ws, Plan& P) {
  // distinct input fields (instances of one source share its pupil field)
  P.field_of.assign(n, -1);
  for (int i = 0; i < n; ++i) {
    int f = -1;
    for (size_t q = 0; q < P.fields.size(); ++q)
      if (P.fields[q] == inst[i].field) { f = (int)q; break; }
    if (f < 0) {
      f = (int)P.fields.size();
      P.fields.push_back(inst[i].field);
      P.field_n.push_back(inst[i].n_in);
    }
    if (P.field_n[f] != inst[i].n_in) {
      set_error("instance %d: the same field pointer is used with n_in %d and %d", i,
                P.field_n[f], inst[i].n_in);
      return DLX_ERR_INVALID;
    }
    P.field_of[i] = f;
  }
  // output planes grouped by output, in instance order within an output
  std::vector<std::vector<int>> members(n_outputs);
  for (int i = 0; i < n; ++i) members[inst[i].out_index].push_back(i);

  size_t off = 0;
  auto take = [&](size_t bytes) { const size_t o = off; off += align_up(bytes); return o; };
  const size_t nf = P.fields.size();
  P.off_cvt = take(nf * sizeof(CvtJob));
  P.off_mat = take(2 * (size_t)n * sizeof(MatJob));
  // each stage's problems: wide list then narrow list (n in total)
  P.off_g1 = take((size_t)n * sizeof(GroupProblem));
  P.off_g2 = take((size_t)n * sizeof(GroupProblem));
  P.off_sum = take((size_t)n_outputs * sizeof(SumJob));
  P.P_off.resize(nf);
  P.off_P = off;
  for (size_t f = 0; f < nf; ++f) P.P_off[f] = take(mat_bytes(P.field_n[f], P.field_n[f], true));
  P.Wx_off.resize(n);
  P.Wy_off.resize(n);
  P.off_W = off;
  for (int i = 0; i < n; ++i) {
    P.Wx_off[i] = take(mat_bytes(inst[i].n_out, inst[i].n_in, true));
    P.Wy_off[i] = inst[i].shift_x == inst[i].shift_y
                      ? P.Wx_off[i]
                      : take(mat_bytes(inst[i].n_out, inst[i].n_in, true));
  }
  P.U_off.resize(n);
  P.off_U = off;
  for (int i = 0; i < n; ++i) P.U_off[i] = take(mat_bytes(inst[i].n_out, inst[i].n_in, true));
  P.I_off.resize(n);
  P.off_I = off;
  for (int o = 0; o < n_outputs; ++o) {
    if (members[o].empty()) continue;
    const size_t npix = (size_t)inst[members[o][0]].n_out * inst[members[o][0]].n_out;
    const size_t base = take(members[o].size() * npix * sizeof(float));
    for (size_t q = 0; q < members[o].size(); ++q)
      P.I_off[members[o][q]] = base + q * npix * sizeof(float);
  }
  P.total = off;
  if (!ws) return DLX_OK;

  // conversion jobs
  long w = 0;
  for (size_t f = 0; f < nf; ++f) {
    CvtJob j{};
    j.in = (const float2*)P.fields[f];
    j.out = ws + P.P_off[f];
    j.rows = P.field_n[f];
    j.k = P.field_n[f];
    j.row_start = w;
    w += j.rows;
    P.cvt.push_back(j);
  }
  P.cvt_work = w;
  // matrix jobs
  w = 0;
  for (int i = 0; i < n; ++i) {
    const dlx_mft_geom g = geom_of(inst[i]);
    for (int axis = 0; axis < 2; ++axis) {
      if (axis == 1 && P.Wy_off[i] == P.Wx_off[i]) continue;
      const double sh = axis == 0 ? inst[i].shift_x : inst[i].shift_y;
      const MftInst p = make_inst(g, inst[i].wavelength, inst[i].pixel_scale_in,
                                  inst[i].pixel_scale_out, sh);
      MatJob j{};
      j.out = ws + (axis == 0 ? P.Wx_off[i] : P.Wy_off[i]);
      j.rows = inst[i].n_out;
      j.cols = inst[i].n_in;
      j.sign_inv = inst[i].inverse ? 1 : 0;
      j.c_scale = p.a_scale; j.c_off = p.a_off;
      j.r_scale = p.b_scale; j.r_off = p.b_off;
      j.row_start = w;
      w += j.rows;
      P.mat.push_back(j);
    }
  }
  P.mat_work = w;
  // GEMM problems
  for (int i = 0; i < n; ++i) {
    const int ni = inst[i].n_in, no = inst[i].n_out;
    const long rbi = split4_row_bytes(ni);
    const double c = mft_norm(geom_of(inst[i]), inst[i].wavelength, inst[i].pixel_scale_in,
                              inst[i].pixel_scale_out);
    P.g1.push_back(make_problem(ws + P.P_off[P.field_of[i]], rbi, ws + P.Wx_off[i], rbi,
                                ws + P.U_off[i], rbi, ni, no, ni, 1.0f));
    P.g2.push_back(make_problem(ws + P.Wy_off[i], rbi, ws + P.U_off[i], rbi,
                                ws + P.I_off[i], no, no, no, ni,
                                (float)(c * std::sqrt(inst[i].weight))));
  }
  split_narrow(P.g1, P.g1n);
  split_narrow(P.g2, P.g2n);
  order_problems(P.g1, P.g1_xcd);
  order_problems(P.g2, P.g2_xcd);
  order_problems(P.g1n, P.g1n_xcd);
  order_problems(P.g2n, P.g2n_xcd);
  // sums
  for (int o = 0; o < n_outputs; ++o) {
    if (members[o].empty()) continue;
    SumJob j{};
    j.out = outs[o];
    j.planes = (const float*)(ws + P.I_off[members[o][0]]);
    j.npix = (long)inst[members[o][0]].n_out * inst[members[o][0]].n_out;
    j.count = (int)members[o].size();
    P.sum.push_back(j);
  }
  return DLX_OK;
}

// ---------------------------------------------------------------------------
// launches, eager or as a replayed graph
// ---------------------------------------------------------------------------
// Every launch reads its jobs and problems from the descriptors at the front of
// the workspace; its grid and by-value arguments depend on the plan's shape
// only.  So one captured graph serves every call with the same workspace and
// the same shape, whatever the fields, wavelengths, weights, shifts, outputs
// and matrix addresses (all uploaded before the launch).
int launch_plan(const Plan& Q, uint8_t* ws, bool with_matrices, hipStream_t s) {
  auto blocks = [](long rows) { return (unsigned)(rows < 16384 ? rows : 16384); };
  if (Q.cvt_work > 0) {
    hipLaunchKernelGGL(group_to_split4_kernel, dim3(blocks(Q.cvt_work)), dim3(256), 0, s,
                       (const CvtJob*)(ws + Q.off_cvt), (int)Q.cvt.size(), Q.cvt_work);
    if (int st = launch_status("group_to_split4_kernel")) return st;
  }
  if (with_matrices && Q.mat_work > 0) {
    hipLaunchKernelGGL(group_dft_split4_kernel, dim3(blocks(Q.mat_work)), dim3(256), 0, s,
                       (const MatJob*)(ws + Q.off_mat), (int)Q.mat.size(), Q.mat_work);
    if (int st = launch_status("group_dft_split4_kernel")) return st;
  }
  auto grouped = [&](size_t off, const std::vector<GroupProblem>& probs, const int (&xcd)[9],
                     int tile_n, int epi) -> int {
    if (probs.empty()) return DLX_OK;
    GroupLaunch l{};
    l.probs = (const GroupProblem*)(ws + off);
    l.nprobs = (int)probs.size();
    l.tile_n = tile_n;
    std::memcpy(l.xcd_start, xcd, sizeof(l.xcd_start));
    DLX_HIP(launch_cgemm_split4_group(l, epi, s));
    return DLX_OK;
  };
  const size_t n1 = Q.g1.size() * sizeof(GroupProblem), n2 = Q.g2.size() * sizeof(GroupProblem);
  if (int st = grouped(Q.off_g1, Q.g1, Q.g1_xcd, 128, EPI_SPLIT_T)) return st;
  if (int st = grouped(Q.off_g1 + n1, Q.g1n, Q.g1n_xcd, 64, EPI_SPLIT_T)) return st;
  if (int st = grouped(Q.off_g2, Q.g2, Q.g2_xcd, 128, EPI_INTENSITY)) return st;
  if (int st = grouped(Q.off_g2 + n2, Q.g2n, Q.g2n_xcd, 64, EPI_INTENSITY)) return st;
  if (Q.sum.empty()) return DLX_OK;
  long most = 0;
  for (const SumJob& j : Q.sum) most = j.npix > most ? j.npix : most;
  const unsigned gx = (unsigned)std::min<long>((most + 255) / 256, 256);
  hipLaunchKernelGGL(group_sum_kernel, dim3(gx, (unsigned)Q.sum.size()), dim3(256), 0, s,
                     (const SumJob*)(ws + Q.off_sum));
  return launch_status("group_sum_kernel");
}

// Everything launch_plan's grids and by-value arguments depend on.
std::vector<long> plan_shape(const Plan& Q, const uint8_t* ws, bool fresh, int dev) {
  std::vector<long> k = {(long)dev, (long)(uintptr_t)ws, (long)fresh, (long)Q.off_mat,
                         (long)Q.mat.size(), Q.mat_work, (long)Q.off_cvt, (long)Q.cvt.size(),
                         Q.cvt_work,
                         (long)Q.off_g1, (long)Q.off_g2, (long)Q.g1.size(), (long)Q.g1n.size(),
                         (long)Q.g2.size(), (long)Q.g2n.size(), (long)Q.off_sum,
                         (long)Q.sum.size()};
  for (const int* x : {Q.g1_xcd, Q.g1n_xcd, Q.g2_xcd, Q.g2n_xcd}) k.insert(k.end(), x, x + 9);
  for (const SumJob& j : Q.sum) k.push_back(j.npix);
  return k;
}

struct GraphEntry {
  hipGraphExec_t exec = nullptr;
  hipEvent_t done = nullptr;     // recorded after the entry's latest launch
  uint64_t tick = 0;
};

struct GraphCache {
  std::mutex mu;
  std::map<std::vector<long>, GraphEntry> map;
  std::map<int, hipStream_t> capture;   // per device: a private stream to capture on
  uint64_t tick = 0, captures = 0, replays = 0;
  int enabled = -1;              // -1: not yet read from the environment
  static constexpr size_t kMax = 64;
};

GraphCache& graphs() {
  static GraphCache c;
  return c;
}

int graphs_on(GraphCache& c) {
  if (c.enabled < 0) {
    // off by default: on ROCm 7.2 a launch of this graph took more host time
    // than its direct launches (50-480 us against 50-380 us per call)
    const char* e = std::getenv("DLUX_AMD_BATCHED_GRAPHS");
    c.enabled = (e && e[0] == '1') ? 1 : 0;
  }
  return c.enabled;
}

void drop_entry(GraphEntry& e) {
  if (e.done) {
    hipEventSynchronize(e.done);   // the exec is destroyed after its last launch ran
    hipEventDestroy(e.done);
  }
  if (e.exec) hipGraphExecDestroy(e.exec);
  e = GraphEntry{};
}

// Launch the plan through the graph cache.  The first sighting of a shape runs
// eagerly (*eager set); the second captures the launches on a private stream
// (capture records and runs nothing), instantiates the graph and launches it
// on `s`; later sightings launch it again.  The caller's stream may be the
// null stream.
int launch_plan_graph(const Plan& Q, uint8_t* ws, bool fresh, hipStream_t s, bool* eager) {
  *eager = true;
  GraphCache& c = graphs();
  std::lock_guard<std::mutex> lock(c.mu);
  if (!graphs_on(c)) return DLX_OK;
  int dev = 0;
  DLX_HIP(hipGetDevice(&dev));
  std::vector<long> key = plan_shape(Q, ws, fresh, dev);
  auto it = c.map.find(key);
  if (it == c.map.end()) {
    if (c.map.size() >= GraphCache::kMax) {
      auto old = c.map.begin();
      for (auto q = c.map.begin(); q != c.map.end(); ++q)
        if (q->second.tick < old->second.tick) old = q;
      drop_entry(old->second);
      c.map.erase(old);
    }
    c.map[key].tick = ++c.tick;
    return DLX_OK;
  }
  GraphEntry& e = it->second;
  e.tick = ++c.tick;
  if (!e.exec) {
    hipStream_t& cs = c.capture[dev];
    if (!cs) DLX_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    hipGraph_t g = nullptr;
    DLX_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
    const int st = launch_plan(Q, ws, fresh, cs);
    const hipError_t end = hipStreamEndCapture(cs, &g);
    if (st) {
      if (end == hipSuccess && g) hipGraphDestroy(g);
      return st;
    }
    DLX_HIP(end);
    const hipError_t inst = hipGraphInstantiate(&e.exec, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (inst != hipSuccess) {
      e.exec = nullptr;
      DLX_HIP(inst);
    }
    DLX_HIP(hipEventCreateWithFlags(&e.done, hipEventDisableTiming));
    ++c.captures;
  }
  DLX_HIP(hipGraphLaunch(e.exec, s));
  DLX_HIP(hipEventRecord(e.done, s));
  ++c.replays;
  *eager = false;
  return DLX_OK;
}

template <typename T>
void put(std::vector<uint8_t>& host, size_t off, const std::vector<T>& v) {
  if (!v.empty()) std::memcpy(host.data() + off, v.data(), v.size() * sizeof(T));
}

}  // namespace
}  // namespace dlx

using namespace dlx;

extern "C" {

size_t dlx_mft_batched_workspace_size(int32_t n, const dlx_mft_instance* instances,
                                      int32_t n_outputs) {
  if (n <= 0 || !instances || n_outputs <= 0) return 0;
  for (int i = 0; i < n; ++i)
    if (instances[i].out_index < 0 || instances[i].out_index >= n_outputs ||
        instances[i].n_in <= 0 || instances[i].n_out <= 0)
      return 0;
  Plan P;
  if (plan_batched(n, instances, n_outputs, nullptr, nullptr, P) != DLX_OK) return 0;
  return P.total;
}

int dlx_mft_batched_psf(int32_t n, const dlx_mft_instance* instances, int32_t n_outputs,
                        float* const* outputs, void* workspace, size_t workspace_bytes,
                        void* stream) {
  clear_error();
  if (int st = validate(n, instances, n_outputs)) return st;
  if (n == 0 || n_outputs == 0) return DLX_OK;
  DLX_CHECK_ARG(outputs != nullptr, "outputs is NULL");
  for (int i = 0; i < n; ++i)
    DLX_CHECK_ARG(outputs[instances[i].out_index] != nullptr, "outputs[%d] is NULL",
                  instances[i].out_index);
  hipStream_t s = (hipStream_t)stream;
  Plan P;
  if (int st = plan_batched(n, instances, n_outputs, nullptr, nullptr, P)) return st;
  if (!workspace || workspace_bytes < P.total) {
    set_error("dlx_mft_batched_psf workspace too small: need %zu bytes, got %zu", P.total,
              workspace_bytes);
    return DLX_ERR_WORKSPACE;
  }
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  DLX_HIP(hipStreamIsCapturing(s, &cs));
  if (cs != hipStreamCaptureStatusNone) {
    set_error("dlx_mft_batched_psf uploads per-call descriptors and cannot be captured");
    return DLX_ERR_UNSUPPORTED;
  }
  Plan Q;
  uint8_t* ws = (uint8_t*)workspace;
  if (int st = plan_batched(n, instances, n_outputs, outputs, ws, Q)) return st;

  // Transfer matrices depend only on the instances' geometry, wavelengths,
  // scales and shifts: keep the whole set resident in the device cache, keyed
  // by every matrix's parameters, and regenerate only on a miss.
  std::vector<double> key;
  key.reserve(Q.mat.size() * 7);
  for (const MatJob& j : Q.mat)
    key.insert(key.end(), {(double)j.rows, (double)j.cols, (double)j.sign_inv, j.c_scale,
                           j.c_off, j.r_scale, j.r_off});
  const size_t wbytes = Q.off_U - Q.off_W;
  void* wcache = nullptr;
  CacheLease whandle(s);
  bool fresh = true;
  if (int st = cache_acquire_raw(key, wbytes, s, &wcache, &whandle.h, &fresh)) return st;
  if (wcache) {
    uint8_t* w0 = ws + Q.off_W;
    auto rebase = [&](const uint8_t* p) -> uint8_t* {
      return (p >= w0 && p < w0 + wbytes) ? (uint8_t*)wcache + (p - w0) : (uint8_t*)p;
    };
    for (MatJob& j : Q.mat) j.out = rebase(j.out);
    for (GroupProblem& g : Q.g1) g.B = rebase(g.B);
    for (GroupProblem& g : Q.g2) g.A = rebase(g.A);
    for (GroupProblem& g : Q.g1n) g.B = rebase(g.B);
    for (GroupProblem& g : Q.g2n) g.A = rebase(g.A);
  }
  // all descriptors sit at the front of the workspace: one upload
  std::vector<uint8_t> host(Q.off_P, 0);
  put(host, Q.off_cvt, Q.cvt);
  put(host, Q.off_mat, Q.mat);
  put(host, Q.off_g1, Q.g1);
  put(host, Q.off_g1 + Q.g1.size() * sizeof(GroupProblem), Q.g1n);
  put(host, Q.off_g2, Q.g2);
  put(host, Q.off_g2 + Q.g2.size() * sizeof(GroupProblem), Q.g2n);
  put(host, Q.off_sum, Q.sum);
  if (int st = upload(host.data(), host.size(), ws, s)) return st;

  // the launches, as a replayed graph when this shape was seen before
  bool eager = true;
  if (cs == hipStreamCaptureStatusNone)
    if (int st = launch_plan_graph(Q, ws, fresh, s, &eager)) return st;
  if (eager)
    if (int st = launch_plan(Q, ws, fresh, s)) return st;
  if (fresh)
    if (int st = cache_ready(whandle.h, s)) return st;
  return whandle.release();
}

int dlx_batched_graph_stats(uint64_t* captures, uint64_t* launches, int32_t* entries) {
  clear_error();
  GraphCache& c = graphs();
  std::lock_guard<std::mutex> lock(c.mu);
  if (captures) *captures = c.captures;
  if (launches) *launches = c.replays;
  if (entries) {
    int n = 0;
    for (auto& kv : c.map) n += kv.second.exec != nullptr;
    *entries = n;
  }
  return DLX_OK;
}

int dlx_batched_graphs(int32_t enable) {
  clear_error();
  GraphCache& c = graphs();
  std::lock_guard<std::mutex> lock(c.mu);
  const int prev = graphs_on(c);
  if (enable >= 0) {
    c.enabled = enable ? 1 : 0;
    if (!c.enabled) {
      for (auto& kv : c.map) drop_entry(kv.second);
      c.map.clear();
    }
  }
  return prev;
}

}  // extern "C"
