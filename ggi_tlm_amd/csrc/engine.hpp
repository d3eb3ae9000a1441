// engine.hpp -- host-side state of one solve (tlm_handle) and the helpers
// shared by the single-solve C ABI (tlmscn.hip) and the batch engine
// (batch.hip).  Internal to libtlmscn.so.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <type_traits>
#include <vector>

#include "scn_kernels.hpp"

// Cells of padding after each link plane (plane stride = storage cells +
// TLM_PLANE_PAD): one 128-byte line keeps the stride off a power of two
// (a 256^3 grid has 2^24 storage cells), which was 1% faster in the plain
// form and neutral in the precision form.
#ifndef TLM_PLANE_PAD
#define TLM_PLANE_PAD 32
#endif

namespace engine {

using scn::Dev;
using scn::PullArgs;

enum class Kind { INCIDENT, SCATTERED };

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
};

}  // namespace engine

struct tlm_handle {
  using Dev = engine::Dev;
  using Kind = engine::Kind;
  using DevBuf = engine::DevBuf;
  tlm_grid g{};
  int form = TLM_FORM_PLAIN_FP64;
  std::string err;
  int nx = 0, ny = 0, nz = 0;
  long long N = 0, P = 0;
  int device = 0;  // the GPU current at tlm_create, which holds the handle's memory
  hipStream_t stream = nullptr;

  // host tables (reference layout)
  bool cells_set = false, faces_set = false;
  std::vector<int32_t> cell_code, c_mat, c_filt, c_pml, c_exc, c_out, c_cable;
  int n_special = 0;
  std::vector<int32_t> face_code, f_mat, f_exc, f_out, f_cable;
  int n_special_faces = 0;
  std::vector<tlm_volume_material> vmat;
  std::vector<int32_t> smat;
  std::vector<tlm_pml_parameter> pml_par;
  std::vector<int32_t> pml_cell_param;
  int n_functions = 0;
  std::vector<double> value, value_face;
  std::vector<tlm_point_source> ps;
  int n_cell_rows = 0, n_face_rows = 0;
  int huy_function = 0;
  double huy_offset_min = 0.0, Ei[3] = {0, 0, 0}, Hi[3] = {0, 0, 0};
  std::vector<tlm_huygens_face> huy;
  std::vector<int32_t> pt_row, pt_comp;
  std::vector<double> freq;

  // device
  bool finalized = false;
  std::vector<DevBuf> bufs;
  // Link buffers.  cur holds the state; prv is the buffer the last step's
  // incident voltages came from (pulled, prv_mode 1, or read as they were,
  // 2; 0 before the first step), which the PML cells pull again to recompute
  // their previous incident voltages.  A step writes the third buffer.
  // Without PML cells two buffers alternate.
  void* planes[3] = {nullptr, nullptr, nullptr};
  int n_bufs = 2;
  int cur = 0, prv = 1;
  int prv_mode = 0;
  bool prv_apply_r = false;
  double prv_small = 0.0;
  Dev d{};
  int n_gen = 0, n_pml = 0, n_mat = 0;
  int n_pure_pml = 0, n_pure_mat = 0, n_pure_gen = 0, n_mixed_blocks = 0;
  int n_free_gen = 0;  // the first n_free_gen generic list entries: code 0 (free space beside a special face)
  int n_pml_fd_pure = 0, n_pml_fd_all = 0;  // PML cells with a special face
  int n_huy_coded = 0;  // descriptor entries in the kFdHuy form
  int n_pm = 0;         // PML cells holding a material
  int stz_n = 0;        // TLM_set_small_to_zero every stz_n steps (0: off)
  double stz_value = 0.0;
  long long n_mixed_cells = 0;
  int mixed_variant = -1;                 // 0: lane slots, 1: LDS-staged, -1: default (LDS-staged)
  // workgroups of the bulk line kernel: 256-cell storage blocks holding a
  // pure bulk line, in storage order (XCD mappings 0, 1) and interleaved by
  // y-band (mapping 2; -1 entries idle)
  std::vector<int32_t> bulk_blocks;
  const int4* d_mat_recs = nullptr;      // k_mat_lines records (two int4 each)
  int n_mat_recs = 0, n_mat_recs_low = 0;  // the first n_mat_recs_low: materials of filter order <= 1
  long long n_matline_cells = 0, n_matline_mat = 0;  // cells of material lines, their material cells
  int64_t matline_filter_bytes = 0;
  const int4* d_bulk_chunks = nullptr;   // k_bulk_chunks records
  int n_bulk_chunks = 0;
  const int4* d_pml_chunks = nullptr;    // k_pml_chunks records (two int4 each): unflagged chunks, then flagged
  int n_pml_chunks = 0, n_pml_chunks_fast = 0;
  const int32_t* d_bulk_blocks = nullptr;
  const int32_t* d_bulk_blocks_yband = nullptr;
  int n_bulk_blocks_yband = 0;
  const int32_t* d_lds_lines = nullptr;   // LDS-staged mixed-line workgroups
  const int32_t* d_lds_fd = nullptr;
  const int32_t* d_lds_slot_p = nullptr;
  const int32_t* d_lds_slot_i = nullptr;
  int n_lds_wg = 0;
  // host-path (cable junction) cells: storage index and 1-based (x,y,z), scan order
  int n_host = 0;
  const int32_t* d_host_cells = nullptr;
  std::vector<int32_t> host_xyz;
  double* d_host_buf = nullptr;   // 12 * n_host staging words
  // host-path faces (thin layer, diode, spice and switch surface materials)
  int n_host_faces = 0;
  std::vector<int32_t> host_face_xyz;  // 4 per face: cx, cy, cz (+ side cell), axis (1-based)
  const int32_t *d_hf_cp = nullptr, *d_hf_cm = nullptr, *d_hf_ax = nullptr;
  double* d_hf_tot = nullptr;     // 4 face voltages per host face
  double* d_hf_buf = nullptr;     // 4 gathered pulses per host face
  bool hf_ready = false;          // the host wrote this step's face voltages
  int n_lds_var[2] = {0, 0};              // LDS workgroups per class-mask variant, in table order
  const int32_t* d_slot_cell = nullptr;  // mixed-line workgroup slots
  const int32_t* d_slot_idx = nullptr;
  const int32_t* d_gen_cell = nullptr;
  const int32_t* d_gen_fd = nullptr;
  const int32_t* d_pml_cell = nullptr;
  const int32_t* d_pml_fd = nullptr;
  int32_t* d_prow = nullptr;
  int32_t* d_pcomp = nullptr;
  double* d_freq = nullptr;
  double* d_dft = nullptr;
  // frequency-domain outputs (tlm_set_frequency_outputs): host copies, the
  // device tables and one complex accumulator per item
  std::vector<tlm_frequency_output> fo;
  std::vector<int32_t> fo_rows, fo_sides, fo_normals;
  const scn::FreqOut* d_fo = nullptr;
  const scn::FreqItem* d_fo_items = nullptr;
  double* d_fo_acc = nullptr;
  int n_fo_items = 0, n_fo_vol = 0, n_fo_surf = 0;
  // time-domain field outputs (tlm_set_field_outputs): per output its
  // items, a device ring of fd_cap frames, the frames in it, the timestep
  // of the oldest unread frame and the frames already moved to the host
  std::vector<tlm_field_output> fd;
  std::vector<int32_t> fd_rows, fd_sides, fd_normals;
  std::vector<const scn::FreqItem*> d_fd_items;
  std::vector<double*> d_fd_ring;
  std::vector<int> fd_cap, fd_used, fd_first, fd_host_frames;
  std::vector<std::vector<double>> fd_host;
  int n_fd_vol = 0, n_fd_surf = 0;
  // far-field surfaces (tlm_set_far_field_surfaces): their J and M
  // accumulators are frequency-output items after the caller's, from item
  // ff_item0 on, six blocks of n_faces per surface
  std::vector<tlm_far_field_surface> ff;
  std::vector<int32_t> ff_rows, ff_sides, ff_normals;
  std::vector<double> ff_centres;
  std::vector<const double*> d_ff_centres;
  double* d_ff_e = nullptr;  // transform output, sized for the largest surface
  int ff_item0 = 0;
  // RCS surface (tlm_set_rcs_surface): faces sorted by K.r, the far-field
  // time series (4 per point)
  bool rcs_on = false;
  tlm_rcs_surface rcs{};
  std::vector<int32_t> rcs_rows, rcs_sides, rcs_normals;
  std::vector<double> rcs_centres;
  std::vector<scn::RcsFace> rcs_faces;  // sorted, host copy
  const scn::RcsFace* d_rcs_faces = nullptr;
  double* d_rcs = nullptr;
  int64_t dev_bytes = 0;

  // bulk schedule (tlm_set_schedule)
  int sched_variant = -1, sched_by = 1, sched_zc = 16, sched_xcd = 0;  // -1: per-form default
  int layout = TLM_LAYOUT_X_SEGMENTS;     // tlm_set_layout
  int streams = 2;                        // tlm_set_streams
  hipStream_t side = nullptr;             // second stream of a step (created on first use)
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;

  // state machine
  Kind kind = Kind::INCIDENT;
  bool pend_connect = false, pend_boundary = false;
  int pend_ts = 0;

  // profiling: event pairs per kernel class, resolved on read
  bool profiling = false;
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev;
  int64_t prof_launches[TLM_N_KCLASS] = {0};
  double prof_ms[TLM_N_KCLASS] = {0};

  // output ring
  int cap = 4096;
  int ring_used = 0;         // slots holding unread steps on the device
  int ring_first_step = 0;   // timestep of slot 0
  std::vector<double> host_out;  // drained rows awaiting read
  int host_first = 0, host_steps = 0;
  // face outputs: one block of n_face_out_rows * 2 * 6 per connect
  int fring_used = 0, fring_first_step = 0;
  std::vector<double> host_fout;
  int host_ffirst = 0, host_fsteps = 0;
};

namespace engine {

inline int fail(tlm_handle* h, int code, const std::string& msg) {
  if (h) h->err = msg;
  return code;
}

inline int hip_fail(tlm_handle* h, hipError_t e, const char* what) {
  return fail(h, TLM_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

inline int hip_ok(tlm_handle* h, hipError_t e) { return e == hipSuccess ? TLM_OK : hip_fail(h, e, "HIP call"); }

#define HIP_TRY(h, expr)                                   \
  do {                                                     \
    hipError_t e_ = (expr);                                \
    if (e_ != hipSuccess) return hip_fail((h), e_, #expr); \
  } while (0)

// The link buffer the next step or connect writes: neither the one it reads
// (cur) nor, with three buffers, the one the PML cells pull again (prv).
inline int out_buffer(const tlm_handle* h) { return h->n_bufs == 3 ? 3 - h->cur - h->prv : h->cur ^ 1; }

template <typename V>
inline int dev_upload(tlm_handle* h, const std::vector<V>& v, const V** out) {
  size_t bytes = std::max<size_t>(v.size(), 1) * sizeof(V);
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, bytes);
  if (e != hipSuccess) return fail(h, TLM_ERR_NOMEM, "hipMalloc failed for a table");
  h->bufs.push_back({p, bytes});
  h->dev_bytes += (int64_t)bytes;
  if (!v.empty()) HIP_TRY(h, hipMemcpy(p, v.data(), v.size() * sizeof(V), hipMemcpyHostToDevice));
  *out = (const V*)p;
  return TLM_OK;
}

inline int dev_alloc(tlm_handle* h, size_t bytes, void** out) {
  void* p = nullptr;
  bytes = std::max<size_t>(bytes, 8);
  hipError_t e = hipMalloc(&p, bytes);
  if (e != hipSuccess) return fail(h, TLM_ERR_NOMEM, "hipMalloc failed (" + std::to_string(bytes) + " bytes)");
  h->bufs.push_back({p, bytes});
  h->dev_bytes += (int64_t)bytes;
  *out = p;
  return TLM_OK;
}

// Element size of a link plane (the 128-byte line holds 128 / link_size
// cells) and the bytes one link takes in HBM (the extended form adds its
// residual byte).
inline size_t link_size(const tlm_handle* h) { return h->form == TLM_FORM_PLAIN_FP64 ? sizeof(double) : sizeof(float); }
inline size_t link_bytes(const tlm_handle* h) { return link_size(h) + (h->form == TLM_FORM_EXTENDED_FP32 ? 1 : 0); }
inline bool valid_form(int form) {
  return form == TLM_FORM_PLAIN_FP64 || form == TLM_FORM_PRECISION_FP32 || form == TLM_FORM_EXTENDED_FP32;
}

// f(T{}, std::integral_constant<bool, PLAIN>{}) with the link storage type
// of the form: double (plain), float (precision) or scn::xlink (extended).
template <typename F>
inline auto with_form(int form, F&& f) {
  if (form == TLM_FORM_PLAIN_FP64) return f(double{}, std::true_type{});
  if (form == TLM_FORM_EXTENDED_FP32) return f(scn::xlink{}, std::false_type{});
  return f(float{}, std::false_type{});
}

// Defined in tlmscn.hip.
int finalize(tlm_handle* h);
int materialize(tlm_handle* h);
int drain_ring(tlm_handle* h);

}  // namespace engine
