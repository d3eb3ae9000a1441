// batch.hip -- many resident solves stepped together (continuous batching).
//
// A tlm_batch holds solves of any grid size, each a configured tlm_handle
// whose device tables and link planes stay resident in HBM.  One step of
// the batch is one launch per kernel class over the union of the solves'
// work: a block table maps every workgroup to (solve, block of that solve),
// and each solve's timestep, link buffer rotation and output slot live in a
// device record, so the K-step launch sequence is captured once in a HIP
// graph and replayed until the set of solves changes.  Solves that finish
// inside a replay idle (their workgroups return at once) until collected.
// The kernels call the same cell-update functions as the single-solve
// path, so a solve gives the same bits alone and in any mix.
//
// Admission is by HBM budget: tlm_batch_add refuses a solve whose device
// footprint would exceed the budget left, and the caller queues it.
#include "engine.hpp"
#include "../../include/tlmbatch.h"

#include <chrono>

using namespace engine;

namespace engine {

// Device record of one resident solve.
struct BatchSolve {
  Dev d;                     // the solve's tables (d.ring: its whole output series)
  void* planes[3];           // link buffers (planes[2] unused without PML cells)
  int nbuf;
  const int32_t* slot_a;     // mixed-line workgroup slots
  const int32_t* slot_b;
  long long n_slots;
  const int32_t* lds_lines;  // LDS-staged mixed-line workgroups
  const int32_t* lds_fd;
  const int32_t* lds_slot_p;
  const int32_t* lds_slot_i;
  const int4* mat_recs;      // material-line records (two int4 each; order <= 1 materials first)
  const int4* pml_chunks;    // PML chunk records (two int4 each; unflagged chunks first)
  const int4* bulk_chunks;   // bulk chunk records
  int n_pure_pml, n_pure_mat, n_pure_gen, n_free_gen, n_pm, n_huy_coded;
  int cur, prv, prv_mode;    // link buffer rotation, as in the single-solve handle
  int ts, first, last;       // next timestep, first and last timestep
  int stz_n;                 // small-to-zero period (0: off)
  double stz_value;
  int np, nf;                // probe DFT
  const int32_t* prow;
  const int32_t* pcomp;
  const double* freq;
  double* dft;
  double dt;
};

}  // namespace engine

struct tlm_batch {
  struct Entry {
    tlm_handle* h = nullptr;
    int first = 0, last = 0, ts = 0;  // host copy of the device timestep
    int cur0 = 0, prv0 = 1, nbuf = 2;
    double* ring = nullptr;   // cell output rows of every step
    double* fring = nullptr;  // face output blocks of every connect the batch pulls
    int64_t bytes = 0;
  };
  int form = TLM_FORM_PRECISION_FP32;
  int64_t budget = 0, resident = 0;
  std::string err;
  hipStream_t stream = nullptr;
  std::vector<Entry> solves;
  engine::BatchSolve* d_solves = nullptr;
  int d_cap = 0;
  // block tables: lines, mixed (lane slots, or LDS-staged without PML
  // cells), LDS-staged with PML cells, pml, material, generic,
  // pml-material, huygens, dft
  // lines (bulk chunk records), mixed, mixed with PML cells, PML chunks
  // without / with flagged cells, material list, generic code-0 cells and
  // the other generic cells, pml-material, huygens, dft, material lines of
  // order <= 1 materials and of the others
  enum { T_LINES, T_MIXED, T_MIXED_P, T_PML, T_PMLG, T_MAT, T_GENF, T_GEN, T_PM, T_HUY, T_DFT, T_MATL, T_MATLH, T_N };
  int2* d_tab[T_N] = {};
  int n_tab[T_N] = {};
  size_t cap_tab[T_N] = {};
  bool dirty = true;
  int graph_steps = 32;
  int mixed = -1;    // mixed-line kernel: 1 LDS-staged, 0 lane slots, -1 per form
  int streams = 2;   // 2: class kernels on a second stream inside each step
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  hipGraphExec_t graph = nullptr;
  int graph_k = 0;
  double last_ms_per_step = 0.0;
};

namespace engine {

int bfail(tlm_batch* b, int code, const std::string& msg) {
  if (b) b->err = msg;
  return code;
}

#define BHIP_TRY(b, expr)                                                              \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess) return bfail((b), TLM_ERR_HIP, std::string(#expr) + ": " + \
                                                             hipGetErrorString(e_));   \
  } while (0)

// Upper bound of a solve's device bytes before its tables are built: two
// link buffers, class tables, PML and filter states, output series.
int64_t estimate_bytes(const tlm_handle* h, int n_steps) {
  const int64_t ls = link_bytes(h);
  int64_t b = 3 * 12 * ((int64_t)h->N + 64 * TLM_MAX_SEGMENTS + TLM_PLANE_PAD) * ls;  // link planes, padding
  b += (int64_t)h->N / 8 + (int64_t)h->N / 8;            // mask, line table
  const int64_t n_pml = (int64_t)h->pml_cell_param.size();
  b += (n_pml + 64) * (8 * (scn::PS_N + 3) + 12);        // PML state (planes padded to 64) and lists
  b += (int64_t)h->n_special * 64;                       // class lists, slots
  b += (int64_t)h->n_special * 2 * 3 * 8 * scn::NW;      // filter states (bound)
  b += (int64_t)h->n_special_faces * 48;
  int64_t out_rows = 0;
  for (int32_t o : h->c_out) out_rows += o != 0;
  b += (int64_t)n_steps * out_rows * 6 * 8;              // output series
  b += (int64_t)n_steps * (int64_t)h->f_out.size() * 12 * 8;  // face output series (bound)
  return b + (1 << 20);
}

// The buffer a step writes (out_buffer of the single-solve handle).
__host__ __device__ inline int batch_out(const BatchSolve& s) {
  return s.nbuf == 3 ? 3 - s.cur - s.prv : s.cur ^ 1;
}

// Buffer rotation after k steps of a solve that entered with (cur0, prv0):
// the first step reads its incident voltages as they are, the later ones pull.
inline void batch_rotation(int nbuf, int cur0, int prv0, int k, int& cur, int& prv, int& mode) {
  BatchSolve s;
  s.nbuf = nbuf;
  s.cur = cur0;
  s.prv = prv0;
  mode = 0;
  for (int j = 0; j < k; j++) {
    const int out = batch_out(s);
    s.prv = s.cur;
    s.cur = out;
    mode = j > 0 ? 1 : 2;
  }
  cur = s.cur;
  prv = s.prv;
}

template <typename T, bool PLAIN>
__device__ inline void batch_pull_args(const BatchSolve& s, PullArgs& a, bool& pull, int& slot) {
  pull = s.ts > s.first;
  a.apply_r = true;
  a.ts_connect = s.ts - 1;
  a.small = (s.stz_n > 0 && s.ts % s.stz_n == 0) ? s.stz_value : 0.0;
  // the pending connect (of step ts-1) records its face outputs in its own
  // slot of the solve's face series
  a.fslot = (s.ts > s.first && s.d.n_face_out_rows) ? s.ts - 1 - s.first : -1;
  a.prev = s.planes[s.prv];
  a.prev_mode = s.prv_mode;
  a.prev_apply_r = true;
  a.prev_small = (s.stz_n > 0 && (s.ts - 1) % s.stz_n == 0) ? s.stz_value : 0.0;
  slot = s.ts - s.first;
}

// Pure bulk lines of a solve, one wavefront per chunk record (k_bulk_chunks).
template <typename T, bool PLAIN>
__global__ __launch_bounds__(scn::kChunk) void kb_lines(const BatchSolve* __restrict__ S,
                                                        const int2* __restrict__ tab) {
  const int2 e = tab[blockIdx.x];
  const BatchSolve& s = S[e.x];
  if (s.ts > s.last) return;
  PullArgs a;
  bool pull;
  int slot;
  batch_pull_args<T, PLAIN>(s, a, pull, slot);
  const T* Lin = (const T*)s.planes[s.cur];
  T* Lout = (T*)s.planes[batch_out(s)];
  const int4 r = s.bulk_chunks[e.y];
  if (pull) scn::bulk_chunk_cell<T, PLAIN, true>(s.d, Lin, Lout, a, r, threadIdx.x);
  else scn::bulk_chunk_cell<T, PLAIN, false>(s.d, Lin, Lout, a, r, threadIdx.x);
}

// PML cells of pure PML lines of a solve, one wavefront per chunk record
// (k_pml_chunks; FAST: the records without flagged cells).
template <typename T, bool PLAIN, bool FAST>
__global__ __launch_bounds__(scn::kChunk) void kb_pml_chunks(const BatchSolve* __restrict__ S,
                                                             const int2* __restrict__ tab) {
  const int2 e = tab[blockIdx.x];
  const BatchSolve& s = S[e.x];
  if (s.ts > s.last) return;
  PullArgs a;
  bool pull;
  int slot;
  batch_pull_args<T, PLAIN>(s, a, pull, slot);
  const T* Lin = (const T*)s.planes[s.cur];
  T* Lout = (T*)s.planes[batch_out(s)];
  const int4 r0 = s.pml_chunks[2 * e.y], r1 = s.pml_chunks[2 * e.y + 1];
  if (pull) scn::pml_chunk_cell<T, PLAIN, true, FAST>(s.d, Lin, Lout, a, slot, r0, r1, threadIdx.x);
  else scn::pml_chunk_cell<T, PLAIN, false, FAST>(s.d, Lin, Lout, a, slot, r0, r1, threadIdx.x);
}

template <typename T, bool PLAIN, bool PULL>
__device__ inline void mixed_lane(const BatchSolve& s, const T* Lin, T* Lout, const PullArgs& a, int slot,
                                  long long si) {
  const Dev& d = s.d;
  const int32_t w = s.slot_a[si];
  if (w < 0) return;
  const long long c = w & scn::kCellMask;
  const int cls = (w >> 28) & 3;
  if (cls == 0) {
    long long cs[scn::kMixedBulkPerLane];
    cs[0] = c;
#pragma unroll
    for (int k = 1; k < scn::kMixedBulkPerLane; k++) cs[k] = s.slot_b[(k - 1) * s.n_slots + si];
    scn::bulk_update_n<T, PLAIN, PULL, scn::kMixedBulkPerLane>(d, Lin, Lout, a, cs);
    return;
  }
  const int i = s.slot_b[si];
  if (cls == 1) scn::pml_update<T, PLAIN, PULL>(d, Lin, Lout, a, slot, i);
  else if (cls == 2) scn::material_update<T, PLAIN, PULL>(d, Lin, Lout, a, i);
  else scn::generic_update<T, PLAIN, PULL>(d, Lin, Lout, a, s.ts, slot, i);
}

template <typename T, bool PLAIN>
__global__ __launch_bounds__(256) void kb_mixed(const BatchSolve* __restrict__ S, const int2* __restrict__ tab) {
  const int2 e = tab[blockIdx.x];
  const BatchSolve& s = S[e.x];
  if (s.ts > s.last) return;
  PullArgs a;
  bool pull;
  int slot;
  batch_pull_args<T, PLAIN>(s, a, pull, slot);
  const T* Lin = (const T*)s.planes[s.cur];
  T* Lout = (T*)s.planes[batch_out(s)];
  const long long si = (long long)e.y * 256 + threadIdx.x;
  if (pull) mixed_lane<T, PLAIN, true>(s, Lin, Lout, a, slot, si);
  else mixed_lane<T, PLAIN, false>(s, Lin, Lout, a, slot, si);
}

template <typename T, bool PLAIN, int CM>
__global__ __launch_bounds__(scn::kLdsCells) void kb_mixed_lds(const BatchSolve* __restrict__ S,
                                                              const int2* __restrict__ tab) {
  __shared__ double sv[12 * scn::kLdsCells];
  const int2 e = tab[blockIdx.x];
  const BatchSolve& s = S[e.x];
  if (s.ts > s.last) return;  // the whole workgroup leaves together
  PullArgs a;
  bool pull;
  int slot;
  batch_pull_args<T, PLAIN>(s, a, pull, slot);
  const T* Lin = (const T*)s.planes[s.cur];
  T* Lout = (T*)s.planes[batch_out(s)];
  if (pull)
    scn::mixed_lds_body<T, PLAIN, true, CM>(s.d, Lin, Lout, a, e.y, s.lds_lines, s.lds_fd, s.lds_slot_p,
                                            s.lds_slot_i, s.ts, slot, sv);
  else
    scn::mixed_lds_body<T, PLAIN, false, CM>(s.d, Lin, Lout, a, e.y, s.lds_lines, s.lds_fd, s.lds_slot_p,
                                             s.lds_slot_i, s.ts, slot, sv);
}

// Material lines of a solve, one wavefront per record (k_mat_lines; LOW:
// the records of order <= 1 materials).
template <typename T, bool PLAIN, bool LOW>
__global__ __launch_bounds__(scn::kChunk) void kb_mat_lines(const BatchSolve* __restrict__ S,
                                                            const int2* __restrict__ tab) {
  const int2 e = tab[blockIdx.x];
  const BatchSolve& s = S[e.x];
  if (s.ts > s.last) return;
  PullArgs a;
  bool pull;
  int slot;
  batch_pull_args<T, PLAIN>(s, a, pull, slot);
  const T* Lin = (const T*)s.planes[s.cur];
  T* Lout = (T*)s.planes[batch_out(s)];
  const int4 r0 = s.mat_recs[2 * e.y], r1 = s.mat_recs[2 * e.y + 1];
  if (pull) scn::mat_line_cell<T, PLAIN, true, LOW>(s.d, Lin, Lout, a, r0, r1, threadIdx.x);
  else scn::mat_line_cell<T, PLAIN, false, LOW>(s.d, Lin, Lout, a, r0, r1, threadIdx.x);
}

// One class list of a solve: 1 material, 2 generic cells other than code 0,
// 3 PML holding a material, 4 generic code-0 cells (free space beside a
// special face).
template <typename T, bool PLAIN, int CLS>
__global__ __launch_bounds__(256) void kb_list(const BatchSolve* __restrict__ S, const int2* __restrict__ tab) {
  const int2 e = tab[blockIdx.x];
  const BatchSolve& s = S[e.x];
  if (s.ts > s.last) return;
  const int i = (CLS == 2 ? s.n_free_gen : 0) + e.y * 256 + threadIdx.x;
  const int n = CLS == 1 ? s.n_pure_mat : CLS == 2 ? s.n_pure_gen : CLS == 4 ? s.n_free_gen : s.n_pm;
  if (i >= n) return;
  PullArgs a;
  bool pull;
  int slot;
  batch_pull_args<T, PLAIN>(s, a, pull, slot);
  const Dev& d = s.d;
  const T* Lin = (const T*)s.planes[s.cur];
  T* Lout = (T*)s.planes[batch_out(s)];
  if (pull) {
    if (CLS == 1) scn::material_update<T, PLAIN, true>(d, Lin, Lout, a, i);
    else if (CLS == 2) scn::generic_update<T, PLAIN, true>(d, Lin, Lout, a, s.ts, slot, i);
    else if (CLS == 4) scn::generic_update<T, PLAIN, true, true>(d, Lin, Lout, a, s.ts, slot, i);
    else scn::pml_material_update<T, PLAIN, true>(d, Lin, Lout, a, slot, i);
  } else {
    if (CLS == 1) scn::material_update<T, PLAIN, false>(d, Lin, Lout, a, i);
    else if (CLS == 2) scn::generic_update<T, PLAIN, false>(d, Lin, Lout, a, s.ts, slot, i);
    else if (CLS == 4) scn::generic_update<T, PLAIN, false, true>(d, Lin, Lout, a, s.ts, slot, i);
    else scn::pml_material_update<T, PLAIN, false>(d, Lin, Lout, a, slot, i);
  }
}

// Per-step Huygens values of the flattened face codes (pending connect = ts-1).
__global__ __launch_bounds__(256) void kb_huygens(const BatchSolve* __restrict__ S, const int2* __restrict__ tab) {
  const int2 e = tab[blockIdx.x];
  const BatchSolve& s = S[e.x];
  if (s.ts > s.last || s.ts <= s.first) return;
  const int f = e.y * 256 + threadIdx.x;
  if (f < s.d.n_huy) s.d.huy_val[f] = scn::huygens_value(s.d, s.ts - 1, s.d.hf[f].offset);
}

// Probe DFT of this step's output row (k_dft's accumulation).
__global__ __launch_bounds__(256) void kb_dft(const BatchSolve* __restrict__ S, const int2* __restrict__ tab) {
  const int2 e = tab[blockIdx.x];
  const BatchSolve& s = S[e.x];
  if (s.ts > s.last) return;
  const int i = e.y * 256 + threadIdx.x;
  if (i >= s.np * s.nf) return;
  const int p = i / s.nf, k = i - p * s.nf;
  const double v = s.d.ring[((long long)(s.ts - s.first) * s.d.n_out_rows + s.prow[p]) * 6 + s.pcomp[p]];
  const double time = (double)(s.ts - 1) * s.dt;
  const double ph = 2.0 * scn::kPI * s.freq[k] * time;
  double sn, cs;
  sincos(ph, &sn, &cs);
  s.dft[2 * i] += v * cs * s.dt;
  s.dft[2 * i + 1] -= v * sn * s.dt;
}

// End of a batch step: rotate the link buffers and advance the timestep.
__global__ void kb_advance(BatchSolve* S, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  BatchSolve& s = S[i];
  if (s.ts > s.last) return;
  const int out = batch_out(s);
  s.prv_mode = s.ts > s.first ? 1 : 2;
  s.prv = s.cur;
  s.cur = out;
  s.ts += 1;
}

// LDS-staged unless lane slots are asked for (the extended form has the LDS-staged kernel only)
bool batch_lds(const tlm_batch* b) { return b->mixed != 0 || b->form == TLM_FORM_EXTENDED_FP32; }

// One batch step.  With two streams the PML, material, generic and
// PML-material kernels run on the side stream concurrently with the bulk
// and mixed-line kernels: they write disjoint cells and read only the
// previous step's buffers; the probe DFT and the advance follow the join.
// (PML on the side stream and mixed lines on the main one after the bulk
// lines measured 2.7% faster on the serving requests than the other way
// round; the single solve keeps PML on its main stream.)
template <typename T, bool PLAIN>
void launch_batch_step(tlm_batch* b) {
  const hipStream_t st = b->stream;
  const bool fork = b->streams > 1;
  const hipStream_t cs = fork ? b->side : st;
  BatchSolve* S = b->d_solves;
  auto go = [&](hipStream_t s, int t, void (*k)(const BatchSolve*, const int2*), int threads = 256) {
    if (b->n_tab[t]) hipLaunchKernelGGL(k, dim3(b->n_tab[t]), dim3(threads), 0, s, S, b->d_tab[t]);
  };
  go(st, tlm_batch::T_HUY, kb_huygens);
  if (fork) {
    (void)hipEventRecord(b->ev_fork, st);
    (void)hipStreamWaitEvent(b->side, b->ev_fork, 0);
  }
  go(st, tlm_batch::T_LINES, kb_lines<T, PLAIN>, scn::kChunk);
  if (batch_lds(b)) {
    go(st, tlm_batch::T_MIXED, kb_mixed_lds<T, PLAIN, scn::kMixedMasks[0]>, scn::kLdsCells);
    go(st, tlm_batch::T_MIXED_P, kb_mixed_lds<T, PLAIN, scn::kMixedMasks[1]>, scn::kLdsCells);
  } else if constexpr (!std::is_same<T, scn::xlink>::value) {
    go(st, tlm_batch::T_MIXED, kb_mixed<T, PLAIN>);
  }
  go(cs, tlm_batch::T_PML, kb_pml_chunks<T, PLAIN, true>, scn::kChunk);
  go(cs, tlm_batch::T_PMLG, kb_pml_chunks<T, PLAIN, false>, scn::kChunk);
  go(cs, tlm_batch::T_MATL, kb_mat_lines<T, PLAIN, true>, scn::kChunk);
  go(cs, tlm_batch::T_MATLH, kb_mat_lines<T, PLAIN, false>, scn::kChunk);
  go(cs, tlm_batch::T_MAT, kb_list<T, PLAIN, 1>);
  go(cs, tlm_batch::T_GENF, kb_list<T, PLAIN, 4>);
  go(cs, tlm_batch::T_GEN, kb_list<T, PLAIN, 2>);
  go(cs, tlm_batch::T_PM, kb_list<T, PLAIN, 3>);
  if (fork) {
    (void)hipEventRecord(b->ev_join, b->side);
    (void)hipStreamWaitEvent(st, b->ev_join, 0);
  }
  go(st, tlm_batch::T_DFT, kb_dft);
  const int n = (int)b->solves.size();
  hipLaunchKernelGGL(kb_advance, dim3((n + 255) / 256), dim3(256), 0, st, S, n);
}

// Destroy the captured graph.  tlm_batch_run returns while its replays may
// still execute, and a replay reads the graph's kernel nodes and argument
// buffers, the device records and the block tables: the stream is drained
// first, so that neither the graph nor the tables it names are released
// (or, in rebuild, overwritten) while a replay is in flight.
int drop_graph(tlm_batch* b) {
  if (!b->graph) return TLM_OK;
  BHIP_TRY(b, hipStreamSynchronize(b->stream));
  (void)hipGraphExecDestroy(b->graph);
  b->graph = nullptr;
  return TLM_OK;
}

// Rebuild the device records and block tables after the set of solves changed.
int rebuild(tlm_batch* b) {
  int rc;
  if ((rc = drop_graph(b))) return rc;
  const int n = (int)b->solves.size();
  if (n > b->d_cap) {
    if (b->d_solves) BHIP_TRY(b, hipFree(b->d_solves));
    b->d_cap = std::max(n, 2 * b->d_cap);
    BHIP_TRY(b, hipMalloc(&b->d_solves, (size_t)b->d_cap * sizeof(BatchSolve)));
  }
  std::vector<BatchSolve> rec(n);
  std::vector<std::vector<int2>> tab(tlm_batch::T_N);
  for (int i = 0; i < n; i++) {
    const tlm_batch::Entry& e = b->solves[i];
    const tlm_handle* h = e.h;
    BatchSolve& s = rec[i];
    std::memset(&s, 0, sizeof s);
    s.d = h->d;
    s.d.ring = e.ring;
    s.d.face_ring = e.fring;
    s.planes[0] = h->planes[0];
    s.planes[1] = h->planes[1];
    s.planes[2] = h->planes[2];
    s.nbuf = h->n_bufs;
    s.slot_a = h->d_slot_cell;
    s.slot_b = h->d_slot_idx;
    s.n_slots = (long long)h->n_mixed_blocks * 256;
    s.lds_lines = h->d_lds_lines;
    s.lds_fd = h->d_lds_fd;
    s.lds_slot_p = h->d_lds_slot_p;
    s.lds_slot_i = h->d_lds_slot_i;
    s.mat_recs = h->d_mat_recs;
    s.pml_chunks = h->d_pml_chunks;
    s.bulk_chunks = h->d_bulk_chunks;
    s.n_pure_pml = h->n_pure_pml;
    s.n_pure_mat = h->n_pure_mat;
    s.n_pure_gen = h->n_pure_gen;
    s.n_free_gen = h->n_free_gen;
    s.n_pm = h->n_pm;
    s.n_huy_coded = h->n_huy_coded;
    // the host copy of ts / parity is exact: every step advances every
    // unfinished solve by one
    s.ts = e.ts;
    batch_rotation(e.nbuf, e.cur0, e.prv0, e.ts - e.first, s.cur, s.prv, s.prv_mode);
    s.first = e.first;
    s.last = e.last;
    s.stz_n = h->stz_n;
    s.stz_value = h->stz_value;
    s.np = (int)h->pt_row.size();
    s.nf = (int)h->freq.size();
    s.prow = h->d_prow;
    s.pcomp = h->d_pcomp;
    s.freq = h->d_freq;
    s.dft = h->d_dft;
    s.dt = h->g.dt;
    if (e.ts > e.last) continue;  // finished, awaiting collection
    for (int k = 0; k < h->n_bulk_chunks; k++) tab[tlm_batch::T_LINES].push_back(make_int2(i, k));
    if (batch_lds(b)) {
      static const int lt[2] = {tlm_batch::T_MIXED, tlm_batch::T_MIXED_P};
      int wg = 0;
      for (int v = 0; v < 2; v++)
        for (int k = 0; k < h->n_lds_var[v]; k++) tab[lt[v]].push_back(make_int2(i, wg++));
    } else {
      for (int k = 0; k < h->n_mixed_blocks; k++) tab[tlm_batch::T_MIXED].push_back(make_int2(i, k));
    }
    auto add = [&](int t, long long cnt) {
      for (long long k = 0; k < (cnt + 255) / 256; k++) tab[t].push_back(make_int2(i, (int)k));
    };
    for (int k = 0; k < h->n_mat_recs; k++)
      tab[k < h->n_mat_recs_low ? tlm_batch::T_MATL : tlm_batch::T_MATLH].push_back(make_int2(i, k));
    for (int k = 0; k < h->n_pml_chunks; k++)
      tab[k < h->n_pml_chunks_fast ? tlm_batch::T_PML : tlm_batch::T_PMLG].push_back(make_int2(i, k));
    add(tlm_batch::T_MAT, h->n_pure_mat);
    add(tlm_batch::T_GENF, h->n_free_gen);
    add(tlm_batch::T_GEN, h->n_pure_gen - h->n_free_gen);
    add(tlm_batch::T_PM, h->n_pm);
    if (h->n_huy_coded) add(tlm_batch::T_HUY, h->d.n_huy);
    if (s.d.n_out_rows && s.np && s.nf) add(tlm_batch::T_DFT, (long long)s.np * s.nf);
  }
  if (n) BHIP_TRY(b, hipMemcpyAsync(b->d_solves, rec.data(), (size_t)n * sizeof(BatchSolve), hipMemcpyHostToDevice,
                                    b->stream));
  for (int t = 0; t < tlm_batch::T_N; t++) {
    b->n_tab[t] = (int)tab[t].size();
    if (tab[t].size() > b->cap_tab[t]) {
      if (b->d_tab[t]) BHIP_TRY(b, hipFree(b->d_tab[t]));
      b->cap_tab[t] = std::max(tab[t].size(), 2 * b->cap_tab[t]);
      BHIP_TRY(b, hipMalloc(&b->d_tab[t], b->cap_tab[t] * sizeof(int2)));
    }
    if (!tab[t].empty())
      BHIP_TRY(b, hipMemcpyAsync(b->d_tab[t], tab[t].data(), tab[t].size() * sizeof(int2), hipMemcpyHostToDevice,
                                 b->stream));
  }
  BHIP_TRY(b, hipStreamSynchronize(b->stream));
  b->dirty = false;
  return TLM_OK;
}

int capture(tlm_batch* b, int k) {
  int rc;
  if ((rc = drop_graph(b))) return rc;
  hipGraph_t g = nullptr;
  BHIP_TRY(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  for (int i = 0; i < k; i++) {
    with_form(b->form, [&](auto t, auto plain) { launch_batch_step<decltype(t), decltype(plain)::value>(b); });
  }
  hipError_t e = hipStreamEndCapture(b->stream, &g);
  if (e != hipSuccess) return bfail(b, TLM_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
  e = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) return bfail(b, TLM_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(e));
  b->graph_k = k;
  return TLM_OK;
}

int unfinished(const tlm_batch* b) {
  int n = 0;
  for (auto& e : b->solves) n += e.ts <= e.last;
  return n;
}

// Enqueue `steps` steps of every unfinished solve (those that end earlier
// idle); the host copies of the timesteps advance at once.
int run_steps(tlm_batch* b, int steps) {
  int rc;
  if (b->dirty && (rc = rebuild(b))) return rc;
  int left = steps;
  while (left > 0 && unfinished(b)) {
    const int k = std::min(left, b->graph_steps);
    if (k == b->graph_steps) {
      if (!b->graph || b->graph_k != k)
        if ((rc = capture(b, k))) return rc;
      BHIP_TRY(b, hipGraphLaunch(b->graph, b->stream));
    } else {
      for (int i = 0; i < k; i++) {
        with_form(b->form, [&](auto t, auto plain) { launch_batch_step<decltype(t), decltype(plain)::value>(b); });
      }
      BHIP_TRY(b, hipGetLastError());
    }
    for (auto& e : b->solves) e.ts = std::min(e.ts + k, e.last + 1);
    left -= k;
  }
  return TLM_OK;
}

}  // namespace engine

extern "C" {

int tlm_batch_create(int32_t form, int64_t hbm_budget_bytes, tlm_batch** out) {
  if (!out) return TLM_ERR_ARG;
  *out = nullptr;
  if (!valid_form(form)) return TLM_ERR_ARG;
  if (hbm_budget_bytes <= 0) return TLM_ERR_ARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return TLM_ERR_HIP;
  tlm_batch* b = new (std::nothrow) tlm_batch();
  if (!b) return TLM_ERR_NOMEM;
  b->form = form;
  b->budget = hbm_budget_bytes;
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&b->side, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&b->ev_join, hipEventDisableTiming) != hipSuccess) {
    tlm_batch_destroy(b);
    return TLM_ERR_HIP;
  }
  *out = b;
  return TLM_OK;
}

void tlm_batch_destroy(tlm_batch* b) {
  if (!b) return;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  if (b->graph) (void)hipGraphExecDestroy(b->graph);
  for (auto& e : b->solves)
    if (e.ring) (void)hipFree(e.ring);
  for (auto& e : b->solves)
    if (e.fring) (void)hipFree(e.fring);
  if (b->d_solves) (void)hipFree(b->d_solves);
  for (auto* t : b->d_tab)
    if (t) (void)hipFree(t);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  if (b->side) (void)hipStreamDestroy(b->side);
  if (b->ev_fork) (void)hipEventDestroy(b->ev_fork);
  if (b->ev_join) (void)hipEventDestroy(b->ev_join);
  delete b;
}

const char* tlm_batch_last_error(const tlm_batch* b) { return b ? b->err.c_str() : "null batch"; }

int tlm_batch_add(tlm_batch* b, tlm_handle* h, int32_t first_timestep, int32_t n_steps) {
  if (!b || !h) return TLM_ERR_ARG;
  if (h->form != b->form) return bfail(b, TLM_ERR_ARG, "solve and batch numerical forms differ");
  if (n_steps < 1 || first_timestep < 1 || first_timestep + n_steps - 1 > h->g.n_timesteps)
    return bfail(b, TLM_ERR_ARG, "timestep range out of range");
  if (h->kind != Kind::INCIDENT || h->ring_used || h->host_steps || h->fring_used || h->host_fsteps ||
      h->prv_mode != 0)
    return bfail(b, TLM_ERR_STATE, "a batch takes solves that have not been stepped");
  for (auto& e : b->solves)
    if (e.h == h) return bfail(b, TLM_ERR_STATE, "solve is already in the batch");
  const int64_t est = h->finalized ? h->dev_bytes : estimate_bytes(h, n_steps);
  if (b->resident + est > b->budget)
    return bfail(b, TLM_ERR_BUDGET, "HBM budget of the batch is exhausted");
  int rc;
  if ((rc = finalize(h))) return bfail(b, rc, std::string("finalize: ") + h->err);
  if (h->n_host) return bfail(b, TLM_ERR_UNSUPPORTED, "host-path cells are exchanged on the single-solve path");
  if (h->n_host_faces) return bfail(b, TLM_ERR_UNSUPPORTED, "host-path faces are exchanged on the single-solve path");
  if (h->n_fo_items || h->rcs_on || !h->fd.empty())
    return bfail(b, TLM_ERR_UNSUPPORTED,
                 "frequency, far-field, RCS and field outputs are recorded on the single-solve path");
  tlm_batch::Entry e;
  e.h = h;
  e.first = first_timestep;
  e.last = first_timestep + n_steps - 1;
  e.ts = first_timestep;
  e.cur0 = h->cur;
  e.prv0 = h->prv;
  e.nbuf = h->n_bufs;
  const size_t ring_bytes = (size_t)n_steps * h->d.n_out_rows * 6 * sizeof(double);
  if (ring_bytes) {
    void* p = nullptr;
    if (hipMalloc(&p, ring_bytes) != hipSuccess) return bfail(b, TLM_ERR_NOMEM, "output series allocation failed");
    e.ring = (double*)p;
  }
  // face outputs of the connects of steps first .. last-1 (the last one is
  // pulled after collection, on the single-solve path)
  const size_t fring_bytes = (size_t)(n_steps - 1) * h->d.n_face_out_rows * 12 * sizeof(double);
  if (fring_bytes) {
    void* p = nullptr;
    if (hipMalloc(&p, fring_bytes) != hipSuccess) {
      if (e.ring) (void)hipFree(e.ring);
      return bfail(b, TLM_ERR_NOMEM, "face output series allocation failed");
    }
    e.fring = (double*)p;
  }
  // the handle's own stream may still be uploading its tables
  BHIP_TRY(b, hipStreamSynchronize(h->stream));
  e.bytes = h->dev_bytes + (int64_t)ring_bytes + (int64_t)fring_bytes;
  b->resident += e.bytes;
  b->solves.push_back(e);
  b->dirty = true;
  return TLM_OK;
}

int tlm_batch_run(tlm_batch* b, int32_t n_steps) {
  if (!b || n_steps < 0) return TLM_ERR_ARG;
  return run_steps(b, n_steps);
}

int tlm_batch_synchronize(tlm_batch* b) {
  if (!b) return TLM_ERR_ARG;
  BHIP_TRY(b, hipStreamSynchronize(b->stream));
  return TLM_OK;
}

int tlm_batch_collect(tlm_batch* b, tlm_handle** done, int32_t max_done, int32_t* n_done) {
  if (!b || (max_done && !done)) return TLM_ERR_ARG;
  BHIP_TRY(b, hipStreamSynchronize(b->stream));
  int n = 0;
  std::vector<tlm_batch::Entry> keep;
  for (auto& e : b->solves) {
    if (e.ts <= e.last || n >= max_done) {
      keep.push_back(e);
      continue;
    }
    tlm_handle* h = e.h;
    const int steps = e.last - e.first + 1;
    const size_t w = (size_t)h->d.n_out_rows * 6;
    if (w) {
      h->host_out.resize((size_t)steps * w);
      BHIP_TRY(b, hipMemcpy(h->host_out.data(), e.ring, (size_t)steps * w * sizeof(double), hipMemcpyDeviceToHost));
      h->host_first = e.first;
      h->host_steps = steps;
    }
    if (e.ring) BHIP_TRY(b, hipFree(e.ring));
    const size_t fw = (size_t)h->d.n_face_out_rows * 12;
    if (fw && steps > 1) {
      // connects first .. last-1; the last connect's block follows when the
      // handle applies it (tlm_read_face_outputs, tlm_connect, tlm_step)
      h->host_fout.resize((size_t)(steps - 1) * fw);
      BHIP_TRY(b, hipMemcpy(h->host_fout.data(), e.fring, (size_t)(steps - 1) * fw * sizeof(double),
                            hipMemcpyDeviceToHost));
      h->host_ffirst = e.first;
      h->host_fsteps = steps - 1;
    }
    if (e.fring) BHIP_TRY(b, hipFree(e.fring));
    // the handle continues as after tlm_step(first, n): last connect pending
    batch_rotation(e.nbuf, e.cur0, e.prv0, steps, h->cur, h->prv, h->prv_mode);
    h->prv_apply_r = true;
    h->prv_small = (h->stz_n > 0 && e.last % h->stz_n == 0) ? h->stz_value : 0.0;
    h->kind = Kind::SCATTERED;
    h->pend_connect = true;
    h->pend_boundary = true;
    h->pend_ts = e.last;
    b->resident -= e.bytes;
    done[n++] = h;
  }
  if (n) {
    b->solves.swap(keep);
    b->dirty = true;
  }
  if (n_done) *n_done = n;
  return TLM_OK;
}

int tlm_batch_info(const tlm_batch* b, int32_t* n_solves, int32_t* n_unfinished, int64_t* resident_bytes) {
  if (!b) return TLM_ERR_ARG;
  if (n_solves) *n_solves = (int32_t)b->solves.size();
  if (n_unfinished) *n_unfinished = unfinished(b);
  if (resident_bytes) *resident_bytes = b->resident;
  return TLM_OK;
}

int tlm_batch_set_graph_steps(tlm_batch* b, int32_t k) {
  if (!b || k < 1 || k > 4096) return TLM_ERR_ARG;
  b->graph_steps = k;
  return TLM_OK;
}

int tlm_batch_set_kernels(tlm_batch* b, int32_t mixed, int32_t streams) {
  if (!b) return TLM_ERR_ARG;
  if (mixed < -1 || mixed > 1 || (streams != 1 && streams != 2))
    return bfail(b, TLM_ERR_ARG, "mixed must be -1, 0 or 1 and streams 1 or 2");
  if (mixed == 0 && b->form == TLM_FORM_EXTENDED_FP32)
    return bfail(b, TLM_ERR_UNSUPPORTED, "the extended form runs the LDS-staged mixed-line kernel only");
  if (mixed != b->mixed) b->dirty = true;  // the mixed-line block table changes
  b->mixed = mixed;
  if (streams != b->streams) {
    int rc;
    if ((rc = drop_graph(b))) return rc;
  }
  b->streams = streams;
  return TLM_OK;
}

int tlm_batch_autotune(tlm_batch* b, int32_t steps_per_candidate, int32_t* best_k, int32_t* steps_done) {
  if (!b || steps_per_candidate < 1) return TLM_ERR_ARG;
  struct Cand { int k, mixed; };
  static const Cand cand[] = {{4, 1}, {16, 1}, {64, 1}, {4, 0}, {16, 0}, {64, 0}};
  int rc, done = 0, best = 0;
  double best_ms = 1e300;
  for (int c = 0; c < (int)(sizeof cand / sizeof cand[0]); c++) {
    if (!unfinished(b)) break;
    const int k = cand[c].k;
    if (cand[c].mixed == 0 && b->form == TLM_FORM_EXTENDED_FP32) continue;  // LDS-staged kernel only
    if ((rc = tlm_batch_set_kernels(b, cand[c].mixed, b->streams))) return rc;
    b->graph_steps = k;
    const int n = std::max(k, steps_per_candidate / k * k);
    // first replay captures; time the second
    if ((rc = run_steps(b, k))) return rc;
    BHIP_TRY(b, hipStreamSynchronize(b->stream));
    const auto t0 = std::chrono::steady_clock::now();
    if ((rc = run_steps(b, n))) return rc;
    BHIP_TRY(b, hipStreamSynchronize(b->stream));
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() / n;
    done += k + n;
    if (ms < best_ms) {
      best_ms = ms;
      best = c;
    }
  }
  b->graph_steps = cand[best].k;
  if ((rc = tlm_batch_set_kernels(b, cand[best].mixed, b->streams))) return rc;
  b->last_ms_per_step = best_ms;
  if (best_k) *best_k = cand[best].k;
  if (steps_done) *steps_done = done;
  return TLM_OK;
}

}  // extern "C"
