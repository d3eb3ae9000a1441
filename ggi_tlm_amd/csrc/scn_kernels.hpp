// scn_kernels.hpp -- gfx950 kernels of the SCN scatter/connect hot path.
//
// Layout in HBM: twelve structure-of-arrays link planes per buffer,
// plane p holding port p+1 of every cell.  Cells are stored by x-segment:
// the x range is cut into segments [x0_s, x1_s) where the cell classes of
// whole x-planes change (PML slabs, source planes, walls), and segment s
// holds its cells x-fastest with row width w_s = x1_s - x0_s from storage
// index base_s: c = base_s + (x - x0_s) + w_s*(y + ny*z).  A PML slab or a
// source plane thus fills 128-byte lines by itself.  One segment is the
// reference's scan order c = x + nx*(y + ny*z).  A step reads the previous
// step's scattered voltages from one buffer, pulls each cell's incident
// voltages across its six faces (the reference's connect and
// outer_boundary, SRC/TLM_SOLVER/connect.F90:180-214 and
// outer_boundary.F90:71-131), scatters (scatter.F90:216-585) and writes the
// new scattered voltages to another buffer.  Each link therefore crosses
// HBM once per step in each direction; the neighbour reads hit the caches.
// With PML cells three buffers rotate: the one the previous step pulled
// from is kept, and PML cells pull it again for their previous incident
// voltages instead of storing them.
//
// Every 128-byte line of a link plane is written whole by one kernel.  Lines
// holding one class only run in that class's wavefront-uniform kernel:
//   bulk     free-space cells with six ordinary faces (chunk records)
//   pml      PML cells (scatter_PML.F90) from chunk records, state in SoA
//            arrays (FP64, or 48-bit streamed words in the reduced forms)
//   material dispersive/PEC/PMC cells without sources, probes or special
//            faces (material-line records, or a list)
//   generic  point sources, probes and cells touching a special face
//            (PEC/PMC/free-space surface, face sources), the Thevenin path of
//            scatter.F90
// Lines holding several classes run in the mixed-line kernels, which stage
// a workgroup's lines through LDS (or give each lane one special cell or
// several bulk cells); PML cells holding a material have a kernel of their
// own.
//
// Template parameters: T is the link storage type (double: plain form,
// float: precision form, xlink: extended form); PULL selects between
// reading incident voltages directly (state uploaded after a connect) and
// pulling them from the scattered planes.  In the plain form every
// expression keeps the reference's operation order and the file is
// compiled with -ffp-contract=off, so results are bit-identical to the
// reference's source-order FP64 arithmetic.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "../../include/tlmscn.h"

namespace scn {

constexpr double kZ0 = 376.73031346177;   // SRC/CONSTANTS/constants.F90:29
constexpr double kY0 = 1.0 / kZ0;          // constants.F90:31
constexpr double kC0 = 2.99792458e8;       // constants.F90:39
constexpr double kPI = 3.141592653589793;  // constants.F90:27
constexpr int NW = TLM_MAX_FILTER_ORDER + 1;

// 0-based ports (Mesh_Modules.F90:83-94 minus one)
enum {
  Vy_xmin = 0, Vz_xmin, Vy_xmax, Vz_xmax, Vx_ymin, Vz_ymin,
  Vx_ymax, Vz_ymax, Vx_zmin, Vy_zmin, Vx_zmax, Vy_zmax
};

// PML state words per cell (PML_module.F90:117-129), SoA: state[w*n + i].
// A PML cell without special faces streams only the nine words Vxt, Vyt,
// Vzt and the six series terms; it recomputes Ix, Iy, Iz and Vi (or the
// pair sums) from the previous step's incident voltages (pml_node).
// Plain form: the reference's 24 words, Vi(12) kept port by port.
enum {
  PS_Vxt, PS_Vyt, PS_Vzt, PS_Ix, PS_Iy, PS_Iz,
  PS_Vi,  // 12 words
  PS_Vyxt = PS_Vi + 12, PS_Vzxt, PS_Vxyt, PS_Vzyt, PS_Vyzt, PS_Vxzt,
  PS_N
};
// Reduced forms (FP32 links): the previous step's incident voltages Vi enter
// the update only as the six pair sums (Vznx+Vzpx), (Vynx+Vypx),
// (Vxny+Vxpy), (Vzny+Vzpy), (Vynz+Vypz), (Vxnz+Vxpz)
// (scatter_PML.F90:127-141), so a cell with a special face keeps those and
// Ix, Iy, Iz in FP64: PC_N words per cell.
enum {
  PC_Ix, PC_Iy, PC_Iz,
  PC_Szx, PC_Syx, PC_Sxy, PC_Szy, PC_Syz, PC_Sxz,
  PC_N
};
// The nine words every PML cell streams (Vxt, Vyt, Vzt and the six series
// terms), in the reduced forms as 48-bit values: the double rounded to its
// top 48 bits (36 significand bits, relative error <= 2^-37, far below the
// FP32 links' 2^-24), stored as a u32 plane of the top 32 bits and a u16
// plane of the next 16.  SoA: hi[w*ns + i], lo[w*ns + i], ns = n_pml_state.
enum { SW_Vxt, SW_Vyt, SW_Vzt, SW_Vyxt, SW_Vzxt, SW_Vxyt, SW_Vzyt, SW_Vyzt, SW_Vxzt, SW_N };
// Ixt,Iyt,Izt live in a separate array, touched only by output cells.

// PML cells holding a non-dispersive material (scatter_PML_material.F90):
// the PML_data_type words they use plus PML_material_data_type
// (PML_module.F90:131-156), FP64 SoA in both forms.  PML_cell_data%Ix..Iz,
// stored by the reference from stale scatter locals and never read, are
// not kept.
enum {
  PM_Ixt, PM_Iyt, PM_Izt, PM_Vi,  // Vi: 12 words
  PM_Vxt = PM_Vi + 12, PM_Vyt, PM_Vzt, PM_Vyxt, PM_Vzxt, PM_Vxyt, PM_Vzyt, PM_Vyzt, PM_Vxzt,
  PM_Ixt_m1, PM_Iyt_m1, PM_Izt_m1, PM_Vcx_m1, PM_Vcy_m1, PM_Vcz_m1, PM_Vcx, PM_Vcy, PM_Vcz,
  PM_Vi_m1,  // 12 words
  PM_Vxt_m1 = PM_Vi_m1 + 12, PM_Vyt_m1, PM_Vzt_m1,
  PM_Vyxt_m1, PM_Vzxt_m1, PM_Vxyt_m1, PM_Vzyt_m1, PM_Vyzt_m1, PM_Vxzt_m1,
  PM_Vosxr, PM_Vosyr, PM_Voszr, PM_Vosx, PM_Vosy, PM_Vosz, PM_Vosx_m1, PM_Vosy_m1, PM_Vosz_m1,
  PM_Vssxr, PM_Vssyr, PM_Vsszr, PM_Vsx, PM_Vsy, PM_Vsz, PM_Vsx_m1, PM_Vsy_m1, PM_Vsz_m1,
  PM_N
};

// Packed per-cell word of the PML and material class lists.
constexpr int kCellBits = 28;
constexpr int32_t kCellMask = (1 << kCellBits) - 1;
constexpr int32_t kFlagOut = 1 << 28;  // cell has an output row
constexpr int32_t kFlagFd = 1 << 29;   // cell has a special-face descriptor

// Value for key in a sorted (key, value) table, -1 if absent.
__device__ inline int32_t sorted_lookup(const int2* tab, int n, int32_t key) {
  int lo = 0, hi = n - 1;
  while (lo <= hi) {
    int mid = (lo + hi) >> 1;
    int k = tab[mid].x;
    if (k == key) return tab[mid].y;
    if (k < key) lo = mid + 1; else hi = mid - 1;
  }
  return -1;
}

constexpr int kMaxSeg = TLM_MAX_SEGMENTS;
constexpr int kChunk = 64;  // cells of a chunk record (one wavefront)

struct Dev {
  int nx, ny, nz;
  long long N;        // cells
  long long P;        // plane stride (elements), >= seg_base[nseg]
  // x-segments: segment s holds x in [seg_x0[s], seg_x0[s+1]) from storage
  // index seg_base[s]; seg_base[nseg] ends the storage (bases are multiples
  // of 64 cells; the cells between a segment's end and the next base are
  // padding, never read)
  int nseg;
  int seg_x0[kMaxSeg + 1];
  long long seg_base[kMaxSeg + 1];
  double dl, dt, c_red;  // c_red = c0 / reduced_c_factor
  int n_timesteps;
  double R[6];        // xmin,xmax,ymin,ymax,zmin,zmax
  const uint32_t* mask;  // 1 = cell not in a pure bulk line
  int line_len;          // cells per 128-byte line of a plane
  const uint32_t* line_info;  // per line: 0 bulk, 1 one special class, 2 mixed
  // pml class
  int n_pml;
  const int32_t* pml_cell;       // cell | kFlagOut | kFlagFd
  const uint16_t* pml_par;       // 0-based parameter row
  const int2* pml_out_tab;       // (list position, output row), sorted
  int n_pml_out;
  const int2* pml_fd_tab;        // (list position, face descriptor), sorted
  int n_pml_fd;
  const tlm_pml_parameter* pml_params;
  double* pml_state;             // PS_N (plain) or PC_N (reduced forms) words * n_pml_state
  uint32_t* pml_hi;              // reduced forms: SW_N top-32-bit planes * n_pml_state
  uint16_t* pml_lo;              // reduced forms: SW_N next-16-bit planes * n_pml_state
  double* pml_It;                // 3 * n_pml_state (Ixt,Iyt,Izt)
  // PML cells holding a material (storage order)
  int n_pm;
  const int32_t* pm_cell;        // cell index
  const int32_t* pm_par;         // PML parameter row (0-based)
  const int32_t* pm_mat;         // volume material number (1-based)
  const int32_t* pm_out;         // output row (0-based) or -1
  const int32_t* pm_fd;          // face descriptor or -1
  double* pm_state;              // PM_N words * n_pm
  long long n_pml_state;
  // material class
  int n_mat;
  const int32_t* mat_cell;
  const uint16_t* mat_num;   // material number (1-based)
  const int32_t* mat_filt;   // filter number (1-based), 0 when not dispersive
  // generic class
  int n_gen;
  const int32_t* gen_cell;
  const int32_t* gen_code;   // reference code (0 = free cell beside a special face)
  const int32_t* gen_mat;    // material number (1-based, 0 = none)
  const int32_t* gen_filt;   // filter number (1-based)
  const int32_t* gen_exc;    // cell excitation row (0-based) or -1
  const int32_t* gen_out;    // output row (0-based) or -1
  const int32_t* gen_fd;     // face descriptor or -1
  const tlm_volume_material* vmat;
  double* eps_w;             // NW * n_filt, SoA by word
  double* mu_w;
  long long n_filt;
  // special faces
  const int32_t* fd;         // 6 special-face codes per descriptor (0 = ordinary; kFdHuy form below)
  const int32_t* sf_mtype;   // per special face code (1-based index): surface material type
  const int32_t* sf_exc;     // excitation row (0-based) or -1
  const int32_t* sf_out;     // face output row (0-based) or -1
  const int32_t* sf_host;    // host-path face number (0-based) or -1
  const double* host_tot;    // 4 face voltages per host-path face, written by the host
  double* face_ring;         // slots of n_face_out_rows * 2 sides * 6 fields
  int n_face_out_rows;
  // excitation
  const double* value;       // (n_timesteps, n_functions) column-major
  const double* value_face;
  const int32_t* cex_start;  // CSR over cell excitation rows -> point sources
  const int32_t* cex_src;
  const int32_t* cex_type;   // 6 per cell row
  const int32_t* fex_start;  // CSR over face excitation rows -> point sources
  const int32_t* fex_src;
  const int32_t* fex_type;   // 6 per face row (side 1 types)
  const int32_t* fhy_start;  // CSR over face excitation rows -> Huygens faces
  const int32_t* fhy_face;
  const tlm_point_source* ps;
  const tlm_huygens_face* hf;
  int huy_function;
  double huy_offset_min;
  double Ei[3], Hi[3];
  double* huy_val;           // per Huygens face: its value at the pending connect's timestep
  int n_huy;
  // outputs
  int n_out_rows;
  double* ring;              // slots of n_out_rows*6
};

// ---------------------------------------------------------------------------
// storage layout (x-segments)
// ---------------------------------------------------------------------------

__device__ inline int seg_width(const Dev& d, int s) { return d.seg_x0[s + 1] - d.seg_x0[s]; }

// Coordinates and segment of storage index c (0 <= c < seg_base[nseg]);
// false for the padding after a segment's cells.
__device__ inline bool cell_coords(const Dev& d, long long c, int& x, int& y, int& z, int& s) {
  s = 0;
  while (s + 1 < d.nseg && c >= d.seg_base[s + 1]) s++;
  const unsigned w = (unsigned)seg_width(d, s);
  const unsigned r = (unsigned)(c - d.seg_base[s]);
  const unsigned row = r / w;
  x = d.seg_x0[s] + (int)(r - row * w);
  z = (int)(row / (unsigned)d.ny);
  y = (int)row - z * d.ny;
  return z < d.nz;
}

// Storage index of cell (x, y, z); s receives its segment.
__device__ inline long long storage_index(const Dev& d, int x, int y, int z, int& s) {
  s = 0;
  while (s + 1 < d.nseg && x >= d.seg_x0[s + 1]) s++;
  return d.seg_base[s] + (x - d.seg_x0[s]) + (long long)seg_width(d, s) * (y + (long long)d.ny * z);
}

// Storage index of the x-1 neighbour of cell c = (x, y, z) in segment s
// (x > 0): the last cell of the same row of segment s-1 at a segment edge.
__device__ inline long long nb_xm(const Dev& d, long long c, int x, int y, int z, int s) {
  if (x > d.seg_x0[s]) return c - 1;
  const long long w = seg_width(d, s - 1);
  return d.seg_base[s - 1] + w * (1 + y + (long long)d.ny * z) - 1;
}

// Storage index of the x+1 neighbour (x < nx-1): the first cell of the same
// row of segment s+1 at a segment edge.
__device__ inline long long nb_xp(const Dev& d, long long c, int x, int y, int z, int s) {
  if (x + 1 < d.seg_x0[s + 1]) return c + 1;
  const long long w = seg_width(d, s + 1);
  return d.seg_base[s + 1] + w * (y + (long long)d.ny * z);
}

// ---------------------------------------------------------------------------
// excitation values on the fly
// ---------------------------------------------------------------------------

// get_interpolated_excitation_value, SRC/EXCITATION/excitation.F90:645-690.
// NINT rounds half away from zero, as round() does.
__device__ inline double huygens_value(const Dev& d, int timestep, double offset) {
  double dt = d.dt;
  double t_offset = (offset - d.huy_offset_min) / d.c_red;
  double local_time = ((double)(timestep - 1)) * dt - t_offset;
  int t1 = (int)round(local_time / dt);
  int t2 = t1 + 1;
  double delta_t = local_time - (double)t1 * dt;
  const double* vf = d.value_face + (long long)(d.huy_function - 1) * d.n_timesteps;
  double v1 = (t1 >= 1 && t1 <= d.n_timesteps) ? vf[t1 - 1] : 0.0;
  double v2 = (t2 >= 1 && t2 <= d.n_timesteps) ? vf[t2 - 1] : 0.0;
  return v1 + delta_t * (v2 - v1) / dt;
}

// A face whose only excitation is one Huygens face with an axis-aligned
// unit normal, on a free-space (or no) surface material, is coded
//   kFdHuy | face << 3 | normal,  normal = 2*axis + (1 if negative)
// in the descriptor table, and its value is read from huy_val (filled once
// per step by k_huygens_values) instead of walking the face row's lists.
constexpr int32_t kFdHuy = 1 << 30;
constexpr int kFdHuyFaceBits = 27;

// Per-step Huygens face values: huy_val[f] = huygens_value(timestep, offset_f).
static __global__ void k_huygens_values(Dev d, int timestep) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f < d.n_huy) d.huy_val[f] = huygens_value(d, timestep, d.hf[f].offset);
}

// face_excitation for a kFdHuy face: the Huygens branch of face_excitation
// below with the face's value and normal taken from the code.
__device__ inline void face_excitation_huy(const Dev& d, int32_t code, double f[12]) {
  for (int k = 0; k < 12; k++) f[k] = 0.0;
  const double value = d.huy_val[(code >> 3) & ((1 << kFdHuyFaceBits) - 1)];
  const int axis = (code & 7) >> 1;
  const double sg = (code & 1) ? -1.0 : 1.0;
  const double nx = axis == 0 ? sg : 0.0, ny = axis == 1 ? sg : 0.0, nz = axis == 2 ? sg : 0.0;
  double Js0 = (ny * value * (d.Hi[2]) - nz * value * (d.Hi[1]));
  double Js1 = (nz * value * (d.Hi[0]) - nx * value * (d.Hi[2]));
  double Js2 = (nx * value * (d.Hi[1]) - ny * value * (d.Hi[0]));
  double Ms0 = -(ny * value * (d.Ei[2]) - nz * value * (d.Ei[1]));
  double Ms1 = -(nz * value * (d.Ei[0]) - nx * value * (d.Ei[2]));
  double Ms2 = -(nx * value * (d.Ei[1]) - ny * value * (d.Ei[0]));
  f[6] = f[6] + Js0; f[7] = f[7] + Js1; f[8] = f[8] + Js2;
  f[9] = f[9] + Ms0; f[10] = f[10] + Ms1; f[11] = f[11] + Ms2;
}

// face_excitation_field(row, side, 1:12) at `timestep`: zero, then the point
// sources in order (excitation_points.F90:231-335), then the Huygens faces
// (huygens_surface.F90:760-842).  Every contribution lands on both sides,
// so one 12-vector serves both.
__device__ inline void face_excitation(const Dev& d, int row, int timestep, double f[12]) {
  for (int k = 0; k < 12; k++) f[k] = 0.0;
  for (int s = d.fex_start[row]; s < d.fex_start[row + 1]; s++) {
    const tlm_point_source& p = d.ps[d.fex_src[s]];
    double v = d.value_face[(long long)(p.function - 1) * d.n_timesteps + (timestep - 1)];
    f[p.component - 1] = f[p.component - 1] + v;
  }
  for (int s = d.fhy_start[row]; s < d.fhy_start[row + 1]; s++) {
    const tlm_huygens_face& h = d.hf[d.fhy_face[s]];
    double value = huygens_value(d, timestep, h.offset);
    double nx = (double)h.nx, ny = (double)h.ny, nz = (double)h.nz;
    double Js0 = (ny * value * (d.Hi[2]) - nz * value * (d.Hi[1]));
    double Js1 = (nz * value * (d.Hi[0]) - nx * value * (d.Hi[2]));
    double Js2 = (nx * value * (d.Hi[1]) - ny * value * (d.Hi[0]));
    double Ms0 = -(ny * value * (d.Ei[2]) - nz * value * (d.Ei[1]));
    double Ms1 = -(nz * value * (d.Ei[0]) - nx * value * (d.Ei[2]));
    double Ms2 = -(nx * value * (d.Ei[1]) - ny * value * (d.Ei[0]));
    f[6] = f[6] + Js0; f[7] = f[7] + Js1; f[8] = f[8] + Js2;
    f[9] = f[9] + Ms0; f[10] = f[10] + Ms1; f[11] = f[11] + Ms2;
  }
}

// ---------------------------------------------------------------------------
// incident voltages (connect + outer boundary of the previous step)
// ---------------------------------------------------------------------------

// Extended link storage (TLM_FORM_EXTENDED_FP32): a buffer holds twelve
// FP32 planes followed by twelve int8 planes.  Byte r of value i refines its
// FP32 value h: V = h + r * u(h), u(h) = ulp(h) / 254.  h is V rounded to
// nearest, so |V - h| <= ulp(h)/2 = 127 u(h) and r = rint((V - h)/u) never
// leaves [-127, 127]; the stored value is within u/2 = ulp(h)/508 of V,
// about 2^-32 relative (FP32 alone: 2^-24), in 5 bytes.  Every decoder
// evaluates the same expression, so a stored link reads back with the same
// bits in every kernel.  xlink has no arithmetic and no conversions: every
// access to such a buffer goes through lk_get / lk_put, with the buffer's
// base pointer and plane stride P.
struct xlink {
  float hi;
};
static_assert(sizeof(xlink) == 4, "xlink is the FP32 plane element");

// Register type of a loaded link value: the storage type itself for FP64
// and FP32 planes, FP64 for the extended form.
template <typename T> struct LinkReg { typedef T type; };
template <> struct LinkReg<xlink> { typedef double type; };

// ulp(h) = 2^(e - 150) for the biased FP32 exponent e of h (e = 1 for zero
// and subnormals, whose spacing is that of the smallest normal binade).
__device__ inline int xl_exp(float h) {
  const int e = (int)((__float_as_uint(h) >> 23) & 0xffu);
  return e ? e : 1;
}
__device__ inline double xl_unit(float h) {
  return __longlong_as_double((long long)(xl_exp(h) + 873) << 52) * (1.0 / 254.0);
}

// Split v into its nearest FP32 value and the residual byte (v - h is exact:
// h is v rounded once; (v - h) * 254 / ulp(h) is exact, a product by 254 and
// a power of two).
__device__ inline void xl_split(double v, float& h, int8_t& r) {
  h = (float)v;
  const double inv = __longlong_as_double((long long)(1173 - xl_exp(h)) << 52) * 254.0;  // 1/u(h)
  r = (int8_t)rint((v - (double)h) * inv);
}

// Residual byte of port p of cell c: the residual planes group four ports
// per 32-bit word, word w = p / 4 of cell c at word index w * P + c, so a
// cell's x-, y- and z-facing port residuals are one word each (loads and
// stores of whole words; 64 lanes write whole 128-byte lines).
__device__ inline long long xl_byte(long long P, int p, long long c) {
  return (((long long)(p >> 2) * P + c) << 2) + (p & 3);
}

// Link value of port p of cell c.
template <typename T>
__device__ inline typename LinkReg<T>::type lk_get(const T* __restrict__ L, long long P, int p, long long c) {
  return L[(long long)p * P + c];
}
__device__ inline double lk_get(const xlink* __restrict__ L, long long P, int p, long long c) {
  const float h = reinterpret_cast<const float*>(L)[(long long)p * P + c];
  const int8_t r = reinterpret_cast<const int8_t*>(L + 12 * P)[xl_byte(P, p, c)];
  return (double)h + (double)r * xl_unit(h);
}

// Link value of port p of cell c0 + k, with c0 the same in every lane of the
// wavefront and k the lane's own offset (0 <= k < 2^28): the address is a
// scalar base plus one 32-bit offset register that every such access of the
// lane shares, instead of a 64-bit address per access.
template <typename T>
__device__ inline typename LinkReg<T>::type lk_get_u(const T* __restrict__ L, long long P, int p, long long c0,
                                                     uint32_t k) {
  return *(const T*)((const char*)(L + ((long long)p * P + c0)) + k * (uint32_t)sizeof(T));
}
__device__ inline double lk_get_u(const xlink* __restrict__ L, long long P, int p, long long c0, uint32_t k) {
  const float h = *(const float*)((const char*)(reinterpret_cast<const float*>(L) + ((long long)p * P + c0)) + k * 4u);
  const int8_t r = *(const int8_t*)(reinterpret_cast<const int8_t*>(L + 12 * P) + xl_byte(P, p, c0) + k * 4u);
  return (double)h + (double)r * xl_unit(h);
}

// Stores of scattered values (non-temporal: read again only next step) and
// plain stores (uploads).
template <typename T>
__device__ inline void lk_put_nt(T* __restrict__ L, long long P, int p, long long c, double v) {
  __builtin_nontemporal_store((T)v, &L[(long long)p * P + c]);
}
__device__ inline void lk_put_nt(xlink* __restrict__ L, long long P, int p, long long c, double v) {
  float h;
  int8_t r;
  xl_split(v, h, r);
  __builtin_nontemporal_store(h, reinterpret_cast<float*>(L) + (long long)p * P + c);
  __builtin_nontemporal_store(r, reinterpret_cast<int8_t*>(L + 12 * P) + xl_byte(P, p, c));
}
template <typename T>
__device__ inline void lk_put(T* __restrict__ L, long long P, int p, long long c, double v) {
  L[(long long)p * P + c] = (T)v;
}
__device__ inline void lk_put(xlink* __restrict__ L, long long P, int p, long long c, double v) {
  float h;
  int8_t r;
  xl_split(v, h, r);
  reinterpret_cast<float*>(L)[(long long)p * P + c] = h;
  reinterpret_cast<int8_t*>(L + 12 * P)[xl_byte(P, p, c)] = r;
}

// All twelve link values of cell c, stored non-temporally.  The extended
// form assembles each residual word in registers and stores it whole.
template <typename T>
__device__ inline void store_links(T* __restrict__ L, long long P, long long c, const double v[12]) {
#pragma unroll
  for (int p = 0; p < 12; p++) lk_put_nt(L, P, p, c, v[p]);
}
__device__ inline void store_links(xlink* __restrict__ L, long long P, long long c, const double v[12]) {
  uint32_t w[3] = {0u, 0u, 0u};
#pragma unroll
  for (int p = 0; p < 12; p++) {
    float h;
    int8_t r;
    xl_split(v[p], h, r);
    __builtin_nontemporal_store(h, reinterpret_cast<float*>(L) + (long long)p * P + c);
    w[p >> 2] |= (uint32_t)(uint8_t)r << (8 * (p & 3));
  }
  uint32_t* R = reinterpret_cast<uint32_t*>(L + 12 * P);
#pragma unroll
  for (int k = 0; k < 3; k++) __builtin_nontemporal_store(w[k], R + (long long)k * P + c);
}

// store_links for cell c0 + k (c0 the same in every lane, as lk_get_u).
template <typename T>
__device__ inline void store_links_u(T* __restrict__ L, long long P, long long c0, uint32_t k, const double v[12]) {
#pragma unroll
  for (int p = 0; p < 12; p++)
    __builtin_nontemporal_store((T)v[p], (T*)((char*)(L + ((long long)p * P + c0)) + k * (uint32_t)sizeof(T)));
}
__device__ inline void store_links_u(xlink* __restrict__ L, long long P, long long c0, uint32_t k,
                                     const double v[12]) {
  uint32_t w[3] = {0u, 0u, 0u};
#pragma unroll
  for (int p = 0; p < 12; p++) {
    float h;
    int8_t r;
    xl_split(v[p], h, r);
    __builtin_nontemporal_store(h, (float*)((char*)(reinterpret_cast<float*>(L) + ((long long)p * P + c0)) + k * 4u));
    w[p >> 2] |= (uint32_t)(uint8_t)r << (8 * (p & 3));
  }
  uint32_t* R = reinterpret_cast<uint32_t*>(L + 12 * P);
#pragma unroll
  for (int j = 0; j < 3; j++)
    __builtin_nontemporal_store(w[j], (uint32_t*)((char*)(R + ((long long)j * P + c0)) + k * 4u));
}

template <typename T>
__device__ inline double ld(const T* __restrict__ L, long long P, int port, long long c) {
  return (double)lk_get(L, P, port, c);
}

// Ordinary face: S = V_min + V_max; incident = S - own
// (connect_normal_to_x.F90:4-10).  The precision form takes the exact
// swap, incident = neighbour, which has no rounding.
template <bool PLAIN, typename A>
__device__ inline A free_face(A own, A nb) {
  if (PLAIN) {
    A S = own + nb;
    return S - own;
  }
  return nb;
}

// Special face (connect_normal_to_{x,y,z}.F90:12-314) seen from one side.
// axis 0/1/2; is_min: the cell owns the face as its min face (ports
// V?_xmin of cell cx in the reference's naming), else it is the max-side
// cell cx-1.  own_a/own_b: this cell's scattered values of the two
// polarisations (a,b) = (y,z) | (x,z) | (x,y); nb_a/nb_b: the other side's.
template <bool PLAIN>
__device__ inline void special_face(const Dev& d, int code, int axis, bool is_min,
                                    double own_a, double own_b, double nb_a, double nb_b,
                                    int ts_connect, int fslot, double& inc_a, double& inc_b) {
  const bool huy = (code & kFdHuy) != 0;
  int mtype = huy ? 0 : d.sf_mtype[code];
  double min_a = is_min ? own_a : nb_a, max_a = is_min ? nb_a : own_a;
  double min_b = is_min ? own_b : nb_b, max_b = is_min ? nb_b : own_b;
  double a_min, a_max, b_min, b_max;
  if (mtype == TLM_SMAT_PEC) {
    a_min = 0.0; a_max = 0.0; b_min = 0.0; b_max = 0.0;
  } else if (mtype == TLM_SMAT_PMC) {
    a_min = 2.0 * min_a; a_max = 2.0 * max_a;
    b_min = 2.0 * min_b; b_max = 2.0 * max_b;
  } else if (mtype == 0 || mtype == TLM_SMAT_FREE_SPACE) {
    a_min = min_a + max_a; a_max = a_min;
    b_min = min_b + max_b; b_max = b_min;
  } else {
    // host-path surface material (thin layer, diode, spice port, switch):
    // the face voltages its host model computed from this face's scattered
    // pulses (connect_normal_to_x.F90:75-283)
    const double* T = d.host_tot + 4 * (long long)d.sf_host[code];
    a_min = T[0]; a_max = T[1];
    b_min = T[2]; b_max = T[3];
  }
  int row = huy ? 0 : d.sf_exc[code];
  if (row >= 0) {
    double f[12];
    if (huy) face_excitation_huy(d, code, f);
    else face_excitation(d, row, ts_connect, f);
    const int32_t* h = d.fex_type + 6 * (long long)row;  // side 1 types for both sides (soft for kFdHuy)
    auto H = [&](int k) { return huy ? 1.0 : (double)h[k]; };
    const double dl = d.dl, Z0 = kZ0;
    const double* Js = f + 6;
    const double* Ms = f + 9;
    if (axis == 0) {
      a_min = a_min * H(1) - f[1] * dl - f[5] * Z0 * dl + Js[1] * dl * Z0 / 2.0 + Ms[2] * dl / 2.0;
      a_max = a_max * H(1) - f[1] * dl + f[5] * Z0 * dl + Js[1] * dl * Z0 / 2.0 - Ms[2] * dl / 2.0;
      b_min = b_min * H(2) - f[2] * dl + f[4] * Z0 * dl + Js[2] * dl * Z0 / 2.0 - Ms[1] * dl / 2.0;
      b_max = b_max * H(2) - f[2] * dl - f[4] * Z0 * dl + Js[2] * dl * Z0 / 2.0 + Ms[1] * dl / 2.0;
    } else if (axis == 1) {
      a_min = a_min * H(0) - f[0] * dl + f[5] * Z0 * dl + Js[0] * dl * Z0 / 2.0 - Ms[2] * dl / 2.0;
      a_max = a_max * H(0) - f[0] * dl - f[5] * Z0 * dl + Js[0] * dl * Z0 / 2.0 + Ms[2] * dl / 2.0;
      b_min = b_min * H(2) - f[2] * dl - f[3] * Z0 * dl + Js[2] * dl * Z0 / 2.0 + Ms[0] * dl / 2.0;
      b_max = b_max * H(2) - f[2] * dl + f[3] * Z0 * dl + Js[2] * dl * Z0 / 2.0 - Ms[0] * dl / 2.0;
    } else {
      a_min = a_min * H(0) - f[0] * dl - f[4] * Z0 * dl + Js[0] * dl * Z0 / 2.0 + Ms[1] * dl / 2.0;
      a_max = a_max * H(0) - f[0] * dl + f[4] * Z0 * dl + Js[0] * dl * Z0 / 2.0 - Ms[1] * dl / 2.0;
      b_min = b_min * H(1) - f[1] * dl + f[3] * Z0 * dl + Js[1] * dl * Z0 / 2.0 - Ms[0] * dl / 2.0;
      b_max = b_max * H(1) - f[1] * dl - f[3] * Z0 * dl + Js[1] * dl * Z0 / 2.0 + Ms[0] * dl / 2.0;
    }
  }
  if (is_min) {
    inc_a = a_min - own_a;
    inc_b = b_min - own_b;
  } else {
    inc_a = a_max - own_a;
    inc_b = b_max - own_b;
  }
  // Face output (connect_normal_to_x.F90:323-361, likewise y, z): this
  // side's row; side 1 (the min cell) pairs the post-connect pulse as
  // "positive" with its pre-connect one, side 2 the other way round.
  if (fslot >= 0 && !huy) {
    const int r = d.sf_out[code];
    if (r >= 0) {
      const double ap = is_min ? inc_a : own_a, an = is_min ? own_a : inc_a;
      const double bp = is_min ? inc_b : own_b, bn = is_min ? own_b : inc_b;
      const double dl = d.dl, Z0 = kZ0;
      double* f = d.face_ring + (((long long)fslot * d.n_face_out_rows + r) * 2 + (is_min ? 0 : 1)) * 6;
      if (axis == 0) {
        f[0] = 0.0;
        f[1] = -(ap + an) / dl;
        f[2] = -(bp + bn) / dl;
        f[3] = 0.0;
        f[4] = (bp - bn) / (Z0 * dl);
        f[5] = -(ap - an) / (Z0 * dl);
      } else if (axis == 1) {
        f[0] = -(ap + an) / dl;
        f[1] = 0.0;
        f[2] = -(bp + bn) / dl;
        f[3] = -(bp - bn) / (Z0 * dl);
        f[4] = 0.0;
        f[5] = (ap - an) / (Z0 * dl);
      } else {
        f[0] = -(ap + an) / dl;
        f[1] = -(bp + bn) / dl;
        f[2] = 0.0;
        f[3] = (bp - bn) / (Z0 * dl);
        f[4] = -(ap - an) / (Z0 * dl);
        f[5] = 0.0;
      }
    }
  }
  (void)PLAIN;
}

struct PullArgs {
  bool apply_r;    // outer_boundary pending
  int ts_connect;  // timestep of the pending connect (face sources)
  double small;    // > 0: incident voltages below it are set to zero first
  int fslot;       // face output ring slot of the pending connect (-1: none)
  // The previous step's incident voltages, which PML cells without special
  // faces recompute instead of keeping them as state: prev_mode 0 none (the
  // state is zero), 1 pulled from the link buffer prev (its outer boundary
  // applied if prev_apply_r), 2 read from prev as they are; prev_small is
  // that step's flush threshold.
  const void* prev = nullptr;
  int prev_mode = 0;
  bool prev_apply_r = false;
  double prev_small = 0.0;
};

// TLM_set_small_to_zero (TLM_SOLVER/TLM_set_small_to_zero.F90:34-68):
// abs(V) < small_to_zero_value -> 0, on the incident voltages of the step
// it is due (run_TLM.F90:236-238).
__device__ inline void flush_small(double v[12], double small) {
  if (small > 0.0) {
#pragma unroll
    for (int p = 0; p < 12; p++)
      if (fabs(v[p]) < small) v[p] = 0.0;
  }
}

// Incident voltages of cell c = (x,y,z) of segment sg in v[12].  fdi: face
// descriptor (-1 = six ordinary faces); FD = false compiles the special-face
// path out for callers whose cells have none.
template <typename T, bool PLAIN, bool FD = true>
__device__ inline void pull_cell(const Dev& d, const T* __restrict__ L, long long c, int x, int y, int z, int sg,
                                 int fdi, const PullArgs& a, double v[12]) {
  const long long P = d.P;
  const long long sy = seg_width(d, sg), sz = sy * d.ny;
  // the cell's own values: the plain form's connect reads all twelve; the
  // reduced forms need a side's two only on the outer boundary or at a
  // special face (an ordinary face takes the neighbour's value as it is)
  double o[12] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  if (PLAIN) {
#pragma unroll
    for (int p = 0; p < 12; p++) o[p] = ld(L, P, p, c);
  }
  int fc[6] = {0, 0, 0, 0, 0, 0};
  if (FD && fdi >= 0) {
#pragma unroll
    for (int k = 0; k < 6; k++) fc[k] = d.fd[6 * fdi + k];
  }
  // per face side: (own ports a,b), (neighbour ports a,b), offset, boundary, R
  const int pa[6] = {Vy_xmin, Vy_xmax, Vx_ymin, Vx_ymax, Vx_zmin, Vx_zmax};
  const int pb[6] = {Vz_xmin, Vz_xmax, Vz_ymin, Vz_ymax, Vy_zmin, Vy_zmax};
  const int qa[6] = {Vy_xmax, Vy_xmin, Vx_ymax, Vx_ymin, Vx_zmax, Vx_zmin};
  const int qb[6] = {Vz_xmax, Vz_xmin, Vz_ymax, Vz_ymin, Vy_zmax, Vy_zmin};
  const bool bnd[6] = {x == 0, x == d.nx - 1, y == 0, y == d.ny - 1, z == 0, z == d.nz - 1};
#pragma unroll
  for (int s = 0; s < 6; s++) {
    double oa = o[pa[s]], ob = o[pb[s]];
    if (!PLAIN && (bnd[s] || (FD && fc[s] != 0))) {
      oa = ld(L, P, pa[s], c);
      ob = ld(L, P, pb[s], c);
    }
    if (bnd[s]) {
      double R = a.apply_r ? d.R[s] : 1.0;
      v[pa[s]] = oa * R;
      v[pb[s]] = ob * R;
      continue;
    }
    const long long cn = s == 0 ? nb_xm(d, c, x, y, z, sg)
                         : s == 1 ? nb_xp(d, c, x, y, z, sg)
                         : s == 2 ? c - sy : s == 3 ? c + sy : s == 4 ? c - sz : c + sz;
    double na = ld(L, P, qa[s], cn);
    double nb = ld(L, P, qb[s], cn);
    if (!FD || fc[s] == 0) {
      v[pa[s]] = free_face<PLAIN>(oa, na);
      v[pb[s]] = free_face<PLAIN>(ob, nb);
    } else {
      double ia, ib;
      special_face<PLAIN>(d, fc[s], s >> 1, (s & 1) == 0, oa, ob, na, nb, a.ts_connect, a.fslot, ia, ib);
      v[pa[s]] = ia;
      v[pb[s]] = ib;
    }
  }
}

// ---------------------------------------------------------------------------
// scatter pieces
// ---------------------------------------------------------------------------

// 12-port scatter (scatter.F90:240-272 and :517-547), in place on v.
__device__ inline void scatter_ports(double v[12], double Vx, double Vy, double Vz,
                                     double Ix, double Iy, double Iz) {
  const double Z0 = kZ0;
  double m, M;
  m = v[Vx_ymin]; M = v[Vx_ymax];
  v[Vx_ymin] = Vx - Z0 * Iz - M;
  v[Vx_ymax] = Vx + Z0 * Iz - m;
  m = v[Vx_zmin]; M = v[Vx_zmax];
  v[Vx_zmin] = Vx + Z0 * Iy - M;
  v[Vx_zmax] = Vx - Z0 * Iy - m;
  m = v[Vy_zmin]; M = v[Vy_zmax];
  v[Vy_zmin] = Vy - Z0 * Ix - M;
  v[Vy_zmax] = Vy + Z0 * Ix - m;
  m = v[Vy_xmin]; M = v[Vy_xmax];
  v[Vy_xmin] = Vy + Z0 * Iz - M;
  v[Vy_xmax] = Vy - Z0 * Iz - m;
  m = v[Vz_xmin]; M = v[Vz_xmax];
  v[Vz_xmin] = Vz - Z0 * Iy - M;
  v[Vz_xmax] = Vz + Z0 * Iy - m;
  m = v[Vz_ymin]; M = v[Vz_ymax];
  v[Vz_ymin] = Vz + Z0 * Ix - M;
  v[Vz_ymax] = Vz - Z0 * Ix - m;
}

// Free-space cell (scatter.F90:222-272).
__device__ inline void scatter_free_plain(double v[12]) {
  const double Z0 = kZ0;
  double Vx = (v[Vx_ymin] + v[Vx_ymax] + v[Vx_zmin] + v[Vx_zmax]) / 2.0;
  double Vy = (v[Vy_xmin] + v[Vy_xmax] + v[Vy_zmin] + v[Vy_zmax]) / 2.0;
  double Vz = (v[Vz_xmin] + v[Vz_xmax] + v[Vz_ymin] + v[Vz_ymax]) / 2.0;
  double Ix = (v[Vz_ymax] - v[Vy_zmax] - v[Vz_ymin] + v[Vy_zmin]) / (2.0 * Z0);
  double Iy = (v[Vx_zmax] - v[Vz_xmax] - v[Vx_zmin] + v[Vz_xmin]) / (2.0 * Z0);
  double Iz = (v[Vy_xmax] - v[Vx_ymax] - v[Vy_xmin] + v[Vx_ymin]) / (2.0 * Z0);
  scatter_ports(v, Vx, Vy, Vz, Ix, Iy, Iz);
}

// Precision form of the same update: Z0*I is the half curl exactly in real
// arithmetic, so the division by 2 Z0 and multiplication by Z0 are folded.
__device__ inline void scatter_free_fast(double v[12]) {
  double Vx = (v[Vx_ymin] + v[Vx_ymax] + v[Vx_zmin] + v[Vx_zmax]) * 0.5;
  double Vy = (v[Vy_xmin] + v[Vy_xmax] + v[Vy_zmin] + v[Vy_zmax]) * 0.5;
  double Vz = (v[Vz_xmin] + v[Vz_xmax] + v[Vz_ymin] + v[Vz_ymax]) * 0.5;
  double Ix = (v[Vz_ymax] - v[Vy_zmax] - v[Vz_ymin] + v[Vy_zmin]) * 0.5;
  double Iy = (v[Vx_zmax] - v[Vz_xmax] - v[Vx_zmin] + v[Vz_xmin]) * 0.5;
  double Iz = (v[Vy_xmax] - v[Vx_ymax] - v[Vy_xmin] + v[Vx_ymin]) * 0.5;
  double m, M;
  m = v[Vx_ymin]; M = v[Vx_ymax];
  v[Vx_ymin] = Vx - Iz - M;  v[Vx_ymax] = Vx + Iz - m;
  m = v[Vx_zmin]; M = v[Vx_zmax];
  v[Vx_zmin] = Vx + Iy - M;  v[Vx_zmax] = Vx - Iy - m;
  m = v[Vy_zmin]; M = v[Vy_zmax];
  v[Vy_zmin] = Vy - Ix - M;  v[Vy_zmax] = Vy + Ix - m;
  m = v[Vy_xmin]; M = v[Vy_xmax];
  v[Vy_xmin] = Vy + Iz - M;  v[Vy_xmax] = Vy - Iz - m;
  m = v[Vz_xmin]; M = v[Vz_xmax];
  v[Vz_xmin] = Vz - Iy - M;  v[Vz_xmax] = Vz + Iy - m;
  m = v[Vz_ymin]; M = v[Vz_ymax];
  v[Vz_ymin] = Vz + Ix - M;  v[Vz_ymax] = Vz - Ix - m;
}

// FP32 arithmetic form of scatter_free_fast, used by the precision-form
// vector bulk kernel (links are stored in FP32 anyway; each output carries
// a few more FP32 roundings than the FP64 evaluation).
__device__ inline void scatter_free_f32(float v[12]) {
  float Vx = (v[Vx_ymin] + v[Vx_ymax] + v[Vx_zmin] + v[Vx_zmax]) * 0.5f;
  float Vy = (v[Vy_xmin] + v[Vy_xmax] + v[Vy_zmin] + v[Vy_zmax]) * 0.5f;
  float Vz = (v[Vz_xmin] + v[Vz_xmax] + v[Vz_ymin] + v[Vz_ymax]) * 0.5f;
  float Ix = (v[Vz_ymax] - v[Vy_zmax] - v[Vz_ymin] + v[Vy_zmin]) * 0.5f;
  float Iy = (v[Vx_zmax] - v[Vz_xmax] - v[Vx_zmin] + v[Vz_xmin]) * 0.5f;
  float Iz = (v[Vy_xmax] - v[Vx_ymax] - v[Vy_xmin] + v[Vx_ymin]) * 0.5f;
  float m, M;
  m = v[Vx_ymin]; M = v[Vx_ymax];
  v[Vx_ymin] = Vx - Iz - M;  v[Vx_ymax] = Vx + Iz - m;
  m = v[Vx_zmin]; M = v[Vx_zmax];
  v[Vx_zmin] = Vx + Iy - M;  v[Vx_zmax] = Vx - Iy - m;
  m = v[Vy_zmin]; M = v[Vy_zmax];
  v[Vy_zmin] = Vy - Ix - M;  v[Vy_zmax] = Vy + Ix - m;
  m = v[Vy_xmin]; M = v[Vy_xmax];
  v[Vy_xmin] = Vy + Iz - M;  v[Vy_xmax] = Vy - Iz - m;
  m = v[Vz_xmin]; M = v[Vz_xmax];
  v[Vz_xmin] = Vz - Iy - M;  v[Vz_xmax] = Vz + Iy - m;
  m = v[Vz_ymin]; M = v[Vz_ymax];
  v[Vz_ymin] = Vz + Ix - M;  v[Vz_ymax] = Vz - Ix - m;
}

// The scattered values are read again only in the next step, after the
// whole grid has streamed through the caches: non-temporal stores keep them
// from displacing the lines this step's neighbours still read (measured 3%
// faster steps in both forms).
template <typename T>
__device__ inline void store_cell(T* __restrict__ L, long long P, long long c, const double v[12]) {
  store_links(L, P, c, v);
}

// evaluate_Zfilter / timeshift_Zfilter (FILTER/filter_subroutines.F90:510-575)
// on SoA state w[k*n + f].
__device__ inline double zf_evaluate(const double* a, const double* b, double* w, long long n,
                                     long long f, int order, double G) {
  double W0 = G;
  for (int i = 1; i <= order; i++) W0 = W0 - b[i] * w[i * n + f];
  W0 = W0 / b[0];
  w[f] = W0;
  double F0 = 0.0;
  for (int i = 0; i <= order; i++) F0 = F0 + a[i] * w[i * n + f];
  return F0;
}
__device__ inline void zf_timeshift(double* w, long long n, long long f, int order) {
  for (int i = order; i >= 1; i--) w[i * n + f] = w[(i - 1) * n + f];
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------

// Tile schedule of the bulk kernels: 64 x by cell tiles of an xy-plane.
// Hardware dispatches workgroup b to XCD b % 8; with xcd_remap the logical
// tile of workgroup b is chosen so that each XCD receives one contiguous
// range of logical tiles (x fastest, then y, then z or z-chunk), so the
// neighbour planes and rows a tile reads were fetched into the same XCD's
// L2 by the tiles just before it.
struct BulkMap {
  int nbx, nby, by;     // tiles in x and y, tile height (rows)
  int nchunks, zc;      // z-chunks of zc planes (zc = 1: one plane per block)
  int nblocks, xcd_remap;
};

__device__ inline void bulk_tile(const BulkMap& m, int& bx, int& byi, int& chunk) {
  int b = blockIdx.x, L = b;
  if (m.xcd_remap) {
    const int per = m.nblocks >> 3;
    L = (b & 7) * per + (b >> 3);
  }
  bx = L % m.nbx;
  const int t = L / m.nbx;
  byi = t % m.nby;
  chunk = t / m.nby;
}

// Incident voltages of a bulk cell from its own scattered values o[] and the
// neighbours' facing ports (x-1, x+1, y-1, y+1, z-1, z+1 pairs).
// A bulk cell's own link values.  The plain form's connect needs all twelve
// ((own + incoming) - own).  The precision form needs them only on the outer
// boundary; the x-facing planes are still read in full, since they prime the
// lines of the x-shifted neighbour reads, but the y- and z-facing planes of
// a row are left to the rows and planes beside it, which read them once.
template <bool PLAIN, typename T, typename O>
__device__ inline void bulk_own(const Dev& d, const T* __restrict__ Lin, long long P, long long c, int y, int z,
                                O o[12]) {
  if (PLAIN) {
#pragma unroll
    for (int p = 0; p < 12; p++) o[p] = (O)lk_get(Lin, P, p, c);
    return;
  }
  o[Vy_xmin] = (O)lk_get(Lin, P, Vy_xmin, c); o[Vz_xmin] = (O)lk_get(Lin, P, Vz_xmin, c);
  o[Vy_xmax] = (O)lk_get(Lin, P, Vy_xmax, c); o[Vz_xmax] = (O)lk_get(Lin, P, Vz_xmax, c);
  o[Vx_ymin] = o[Vz_ymin] = o[Vx_ymax] = o[Vz_ymax] = (O)0;
  o[Vx_zmin] = o[Vy_zmin] = o[Vx_zmax] = o[Vy_zmax] = (O)0;
  if (y == 0) { o[Vx_ymin] = (O)lk_get(Lin, P, Vx_ymin, c); o[Vz_ymin] = (O)lk_get(Lin, P, Vz_ymin, c); }
  if (y == d.ny - 1) { o[Vx_ymax] = (O)lk_get(Lin, P, Vx_ymax, c); o[Vz_ymax] = (O)lk_get(Lin, P, Vz_ymax, c); }
  if (z == 0) { o[Vx_zmin] = (O)lk_get(Lin, P, Vx_zmin, c); o[Vy_zmin] = (O)lk_get(Lin, P, Vy_zmin, c); }
  if (z == d.nz - 1) { o[Vx_zmax] = (O)lk_get(Lin, P, Vx_zmax, c); o[Vy_zmax] = (O)lk_get(Lin, P, Vy_zmax, c); }
}

template <bool PLAIN>
__device__ inline void bulk_incident(const Dev& d, const PullArgs& a, int x, int y, int z, const double o[12],
                                     double xa, double xb, double xc, double xd, double ya, double yb,
                                     double yc, double yd, double za, double zb, double zc_, double zd,
                                     double v[12]) {
  const double Rx0 = a.apply_r ? d.R[0] : 1.0, Rx1 = a.apply_r ? d.R[1] : 1.0;
  const double Ry0 = a.apply_r ? d.R[2] : 1.0, Ry1 = a.apply_r ? d.R[3] : 1.0;
  const double Rz0 = a.apply_r ? d.R[4] : 1.0, Rz1 = a.apply_r ? d.R[5] : 1.0;
  if (x == 0) { v[Vy_xmin] = o[Vy_xmin] * Rx0; v[Vz_xmin] = o[Vz_xmin] * Rx0; }
  else { v[Vy_xmin] = free_face<PLAIN>(o[Vy_xmin], xa); v[Vz_xmin] = free_face<PLAIN>(o[Vz_xmin], xb); }
  if (x == d.nx - 1) { v[Vy_xmax] = o[Vy_xmax] * Rx1; v[Vz_xmax] = o[Vz_xmax] * Rx1; }
  else { v[Vy_xmax] = free_face<PLAIN>(o[Vy_xmax], xc); v[Vz_xmax] = free_face<PLAIN>(o[Vz_xmax], xd); }
  if (y == 0) { v[Vx_ymin] = o[Vx_ymin] * Ry0; v[Vz_ymin] = o[Vz_ymin] * Ry0; }
  else { v[Vx_ymin] = free_face<PLAIN>(o[Vx_ymin], ya); v[Vz_ymin] = free_face<PLAIN>(o[Vz_ymin], yb); }
  if (y == d.ny - 1) { v[Vx_ymax] = o[Vx_ymax] * Ry1; v[Vz_ymax] = o[Vz_ymax] * Ry1; }
  else { v[Vx_ymax] = free_face<PLAIN>(o[Vx_ymax], yc); v[Vz_ymax] = free_face<PLAIN>(o[Vz_ymax], yd); }
  if (z == 0) { v[Vx_zmin] = o[Vx_zmin] * Rz0; v[Vy_zmin] = o[Vy_zmin] * Rz0; }
  else { v[Vx_zmin] = free_face<PLAIN>(o[Vx_zmin], za); v[Vy_zmin] = free_face<PLAIN>(o[Vy_zmin], zb); }
  if (z == d.nz - 1) { v[Vx_zmax] = o[Vx_zmax] * Rz1; v[Vy_zmax] = o[Vy_zmax] * Rz1; }
  else { v[Vx_zmax] = free_face<PLAIN>(o[Vx_zmax], zc_); v[Vy_zmax] = free_face<PLAIN>(o[Vy_zmax], zd); }
}

// Workgroup b runs on XCD b % 8.  With remap, the logical block of b is
// chosen so that each XCD takes one contiguous range of logical blocks
// (needs a block count divisible by 8; otherwise identity).
__device__ inline int xcd_block(int b, int nblocks, int remap) {
  if (!remap || (nblocks & 7)) return b;
  return (b & 7) * (nblocks >> 3) + (b >> 3);
}

// Incident voltages of a cell with six ordinary faces (connect and outer
// boundary), pulled from the link buffer L.  Own values first (bulk_own):
// they stream the cell's x-facing lines in once and the x-shifted neighbour
// reads of the same lines then hit the caches.  (Reading no own values at
// all was measured slower.)
template <typename T, bool PLAIN>
__device__ inline void bulk_pull(const Dev& d, const T* __restrict__ L, const PullArgs& a, long long c, int x, int y,
                                 int z, int sg, double v[12]) {
  const long long P = d.P;
  const long long sy = seg_width(d, sg), sz = sy * d.ny;
  double o[12];
  bulk_own<PLAIN>(d, L, P, c, y, z, o);
  double xa = 0, xb = 0, xc = 0, xd = 0, ya = 0, yb = 0, yc = 0, yd = 0, za = 0, zb = 0, zc_ = 0, zd = 0;
  if (x > 0) {
    const long long n = nb_xm(d, c, x, y, z, sg);
    xa = lk_get(L, P, Vy_xmax, n); xb = lk_get(L, P, Vz_xmax, n);
  }
  if (x + 1 < d.nx) {
    const long long n = nb_xp(d, c, x, y, z, sg);
    xc = lk_get(L, P, Vy_xmin, n); xd = lk_get(L, P, Vz_xmin, n);
  }
  if (y > 0) { ya = lk_get(L, P, Vx_ymax, c - sy); yb = lk_get(L, P, Vz_ymax, c - sy); }
  if (y + 1 < d.ny) { yc = lk_get(L, P, Vx_ymin, c + sy); yd = lk_get(L, P, Vz_ymin, c + sy); }
  if (z > 0) { za = lk_get(L, P, Vx_zmax, c - sz); zb = lk_get(L, P, Vy_zmax, c - sz); }
  if (z + 1 < d.nz) { zc_ = lk_get(L, P, Vx_zmin, c + sz); zd = lk_get(L, P, Vy_zmin, c + sz); }
  bulk_incident<PLAIN>(d, a, x, y, z, o, xa, xb, xc, xd, ya, yb, yc, yd, za, zb, zc_, zd, v);
}

// bulk_pull from two link buffers at once: the current step's incident
// voltages from L (boundary as a says) in v and the previous step's from
// a.prev (boundary as a.prev_apply_r says) in w.  Every load is issued
// before either set is used, so the two pulls cost one memory round trip.
template <typename T, bool PLAIN>
__device__ inline void bulk_pull_pair(const Dev& d, const T* __restrict__ L, const PullArgs& a, long long c, int x,
                                      int y, int z, int sg, double v[12], double w[12]) {
  const T* __restrict__ Q = (const T*)a.prev;
  const long long P = d.P;
  const long long sy = seg_width(d, sg), sz = sy * d.ny;
  typedef typename LinkReg<T>::type R;
  R o[12], q[12];
  bulk_own<PLAIN>(d, L, P, c, y, z, o);
  bulk_own<PLAIN>(d, Q, P, c, y, z, q);
  R m[12], r[12];
#pragma unroll
  for (int k = 0; k < 12; k++) m[k] = r[k] = (R)0;
  if (x > 0) {
    const long long n = nb_xm(d, c, x, y, z, sg);
    m[0] = lk_get(L, P, Vy_xmax, n); m[1] = lk_get(L, P, Vz_xmax, n);
    r[0] = lk_get(Q, P, Vy_xmax, n); r[1] = lk_get(Q, P, Vz_xmax, n);
  }
  if (x + 1 < d.nx) {
    const long long n = nb_xp(d, c, x, y, z, sg);
    m[2] = lk_get(L, P, Vy_xmin, n); m[3] = lk_get(L, P, Vz_xmin, n);
    r[2] = lk_get(Q, P, Vy_xmin, n); r[3] = lk_get(Q, P, Vz_xmin, n);
  }
  if (y > 0) {
    m[4] = lk_get(L, P, Vx_ymax, c - sy); m[5] = lk_get(L, P, Vz_ymax, c - sy);
    r[4] = lk_get(Q, P, Vx_ymax, c - sy); r[5] = lk_get(Q, P, Vz_ymax, c - sy);
  }
  if (y + 1 < d.ny) {
    m[6] = lk_get(L, P, Vx_ymin, c + sy); m[7] = lk_get(L, P, Vz_ymin, c + sy);
    r[6] = lk_get(Q, P, Vx_ymin, c + sy); r[7] = lk_get(Q, P, Vz_ymin, c + sy);
  }
  if (z > 0) {
    m[8] = lk_get(L, P, Vx_zmax, c - sz); m[9] = lk_get(L, P, Vy_zmax, c - sz);
    r[8] = lk_get(Q, P, Vx_zmax, c - sz); r[9] = lk_get(Q, P, Vy_zmax, c - sz);
  }
  if (z + 1 < d.nz) {
    m[10] = lk_get(L, P, Vx_zmin, c + sz); m[11] = lk_get(L, P, Vy_zmin, c + sz);
    r[10] = lk_get(Q, P, Vx_zmin, c + sz); r[11] = lk_get(Q, P, Vy_zmin, c + sz);
  }
  double od[12], qd[12];
#pragma unroll
  for (int k = 0; k < 12; k++) { od[k] = (double)o[k]; qd[k] = (double)q[k]; }
  bulk_incident<PLAIN>(d, a, x, y, z, od, m[0], m[1], m[2], m[3], m[4], m[5], m[6], m[7], m[8], m[9], m[10], m[11], v);
  PullArgs b{a.prev_apply_r, 0, 0.0, -1};
  bulk_incident<PLAIN>(d, b, x, y, z, qd, r[0], r[1], r[2], r[3], r[4], r[5], r[6], r[7], r[8], r[9], r[10], r[11], w);
}

// bulk_pull (PAIR false) or bulk_pull_pair (PAIR true) for lane k of a
// chunk whose first cell c0 is the same in every lane (cell c = c0 + k): the
// own, y-, z- and in-segment x-neighbour values are read at a scalar base
// plus the lane's one offset register; only a lane at a segment's first or
// last column computes its x-neighbour index (nb_xm, nb_xp).  Same loads,
// same values as bulk_pull and bulk_pull_pair.
template <typename T, bool PLAIN, bool PAIR>
__device__ inline void chunk_pull_impl(const Dev& d, const T* __restrict__ L, const PullArgs& a, long long c0,
                                       uint32_t k, int x, int y, int z, int sg, double v[12], double w[12]) {
  const T* __restrict__ Q = PAIR ? (const T*)a.prev : L;
  const long long P = d.P;
  const long long c = c0 + k;
  const long long sy = seg_width(d, sg), sz = sy * d.ny;
  typedef typename LinkReg<T>::type R;
  R o[12], q[12];
  if (PLAIN) {
#pragma unroll
    for (int p = 0; p < 12; p++) {
      o[p] = lk_get_u(L, P, p, c0, k);
      if (PAIR) q[p] = lk_get_u(Q, P, p, c0, k);
    }
  } else {
#pragma unroll
    for (int p = 0; p < 12; p++) o[p] = q[p] = (R)0;
    const int xp[4] = {Vy_xmin, Vz_xmin, Vy_xmax, Vz_xmax};
#pragma unroll
    for (int j = 0; j < 4; j++) {
      o[xp[j]] = lk_get_u(L, P, xp[j], c0, k);
      if (PAIR) q[xp[j]] = lk_get_u(Q, P, xp[j], c0, k);
    }
    if (y == 0) {
      o[Vx_ymin] = lk_get_u(L, P, Vx_ymin, c0, k); o[Vz_ymin] = lk_get_u(L, P, Vz_ymin, c0, k);
      if (PAIR) { q[Vx_ymin] = lk_get_u(Q, P, Vx_ymin, c0, k); q[Vz_ymin] = lk_get_u(Q, P, Vz_ymin, c0, k); }
    }
    if (y == d.ny - 1) {
      o[Vx_ymax] = lk_get_u(L, P, Vx_ymax, c0, k); o[Vz_ymax] = lk_get_u(L, P, Vz_ymax, c0, k);
      if (PAIR) { q[Vx_ymax] = lk_get_u(Q, P, Vx_ymax, c0, k); q[Vz_ymax] = lk_get_u(Q, P, Vz_ymax, c0, k); }
    }
    if (z == 0) {
      o[Vx_zmin] = lk_get_u(L, P, Vx_zmin, c0, k); o[Vy_zmin] = lk_get_u(L, P, Vy_zmin, c0, k);
      if (PAIR) { q[Vx_zmin] = lk_get_u(Q, P, Vx_zmin, c0, k); q[Vy_zmin] = lk_get_u(Q, P, Vy_zmin, c0, k); }
    }
    if (z == d.nz - 1) {
      o[Vx_zmax] = lk_get_u(L, P, Vx_zmax, c0, k); o[Vy_zmax] = lk_get_u(L, P, Vy_zmax, c0, k);
      if (PAIR) { q[Vx_zmax] = lk_get_u(Q, P, Vx_zmax, c0, k); q[Vy_zmax] = lk_get_u(Q, P, Vy_zmax, c0, k); }
    }
  }
  R m[12], r[12];
#pragma unroll
  for (int j = 0; j < 12; j++) m[j] = r[j] = (R)0;
  if (x > 0) {
    if (x > d.seg_x0[sg]) {
      m[0] = lk_get_u(L, P, Vy_xmax, c0 - 1, k); m[1] = lk_get_u(L, P, Vz_xmax, c0 - 1, k);
      if (PAIR) { r[0] = lk_get_u(Q, P, Vy_xmax, c0 - 1, k); r[1] = lk_get_u(Q, P, Vz_xmax, c0 - 1, k); }
    } else {
      const long long n = nb_xm(d, c, x, y, z, sg);
      m[0] = lk_get(L, P, Vy_xmax, n); m[1] = lk_get(L, P, Vz_xmax, n);
      if (PAIR) { r[0] = lk_get(Q, P, Vy_xmax, n); r[1] = lk_get(Q, P, Vz_xmax, n); }
    }
  }
  if (x + 1 < d.nx) {
    if (x + 1 < d.seg_x0[sg + 1]) {
      m[2] = lk_get_u(L, P, Vy_xmin, c0 + 1, k); m[3] = lk_get_u(L, P, Vz_xmin, c0 + 1, k);
      if (PAIR) { r[2] = lk_get_u(Q, P, Vy_xmin, c0 + 1, k); r[3] = lk_get_u(Q, P, Vz_xmin, c0 + 1, k); }
    } else {
      const long long n = nb_xp(d, c, x, y, z, sg);
      m[2] = lk_get(L, P, Vy_xmin, n); m[3] = lk_get(L, P, Vz_xmin, n);
      if (PAIR) { r[2] = lk_get(Q, P, Vy_xmin, n); r[3] = lk_get(Q, P, Vz_xmin, n); }
    }
  }
  if (y > 0) {
    m[4] = lk_get_u(L, P, Vx_ymax, c0 - sy, k); m[5] = lk_get_u(L, P, Vz_ymax, c0 - sy, k);
    if (PAIR) { r[4] = lk_get_u(Q, P, Vx_ymax, c0 - sy, k); r[5] = lk_get_u(Q, P, Vz_ymax, c0 - sy, k); }
  }
  if (y + 1 < d.ny) {
    m[6] = lk_get_u(L, P, Vx_ymin, c0 + sy, k); m[7] = lk_get_u(L, P, Vz_ymin, c0 + sy, k);
    if (PAIR) { r[6] = lk_get_u(Q, P, Vx_ymin, c0 + sy, k); r[7] = lk_get_u(Q, P, Vz_ymin, c0 + sy, k); }
  }
  if (z > 0) {
    m[8] = lk_get_u(L, P, Vx_zmax, c0 - sz, k); m[9] = lk_get_u(L, P, Vy_zmax, c0 - sz, k);
    if (PAIR) { r[8] = lk_get_u(Q, P, Vx_zmax, c0 - sz, k); r[9] = lk_get_u(Q, P, Vy_zmax, c0 - sz, k); }
  }
  if (z + 1 < d.nz) {
    m[10] = lk_get_u(L, P, Vx_zmin, c0 + sz, k); m[11] = lk_get_u(L, P, Vy_zmin, c0 + sz, k);
    if (PAIR) { r[10] = lk_get_u(Q, P, Vx_zmin, c0 + sz, k); r[11] = lk_get_u(Q, P, Vy_zmin, c0 + sz, k); }
  }
  double od[12], qd[12];
#pragma unroll
  for (int j = 0; j < 12; j++) { od[j] = (double)o[j]; qd[j] = (double)q[j]; }
  bulk_incident<PLAIN>(d, a, x, y, z, od, m[0], m[1], m[2], m[3], m[4], m[5], m[6], m[7], m[8], m[9], m[10], m[11], v);
  if (PAIR) {
    PullArgs b{a.prev_apply_r, 0, 0.0, -1};
    bulk_incident<PLAIN>(d, b, x, y, z, qd, r[0], r[1], r[2], r[3], r[4], r[5], r[6], r[7], r[8], r[9], r[10], r[11],
                         w);
  }
}
template <typename T, bool PLAIN>
__device__ inline void chunk_pull_pair(const Dev& d, const T* __restrict__ L, const PullArgs& a, long long c0,
                                       uint32_t k, int x, int y, int z, int sg, double v[12], double w[12]) {
  chunk_pull_impl<T, PLAIN, true>(d, L, a, c0, k, x, y, z, sg, v, w);
}
template <typename T, bool PLAIN>
__device__ inline void chunk_pull(const Dev& d, const T* __restrict__ L, const PullArgs& a, long long c0, uint32_t k,
                                  int x, int y, int z, int sg, double v[12]) {
  chunk_pull_impl<T, PLAIN, false>(d, L, a, c0, k, x, y, z, sg, v, nullptr);
}

// Bulk free-space cell update, one cell per lane: pull the six faces
// (connect and outer boundary) and scatter.
template <typename T, bool PLAIN, bool PULL>
__device__ inline void bulk_update(const Dev& d, const T* __restrict__ Lin, T* __restrict__ Lout,
                                   const PullArgs& a, long long c, int x, int y, int z, int sg) {
  const long long P = d.P;
  double v[12];
  if (PULL) {
    bulk_pull<T, PLAIN>(d, Lin, a, c, x, y, z, sg, v);
  } else {
#pragma unroll
    for (int p = 0; p < 12; p++) v[p] = (double)lk_get(Lin, P, p, c);
  }
  flush_small(v, a.small);
  if (PLAIN) scatter_free_plain(v); else scatter_free_fast(v);
  store_cell(Lout, P, c, v);
}

// The tile kernels below (k_bulk_cell, k_bulk_vec, k_bulk_z) address cells
// as c = x + nx*(y + ny*z): they run on the one-segment (row-major) layout
// only, which the host enforces.

// Bulk free-space cells, one cell per thread (zc = 1): block = 64 x by cells
// of one plane.  Neighbour ports are loaded directly; they are the block's
// own rows or the rows/planes of tiles that ran just before on the same XCD.
template <typename T, bool PLAIN, bool PULL>
__global__ __launch_bounds__(256) void k_bulk_cell(Dev d, const T* __restrict__ Lin, T* __restrict__ Lout,
                                                  PullArgs a, BulkMap m) {
  int bx, byi, z;
  bulk_tile(m, bx, byi, z);
  const int x = bx * 64 + threadIdx.x;
  const int y = byi * m.by + threadIdx.y;
  if (x >= d.nx || y >= d.ny || z >= d.nz) return;
  const long long c = x + (long long)d.nx * (y + (long long)d.ny * z);
  if ((d.mask[c >> 5] >> (c & 31)) & 1u) return;
  bulk_update<T, PLAIN, PULL>(d, Lin, Lout, a, c, x, y, z, 0);
}

// Vector access of W consecutive link values (W*sizeof(T) = 16 bytes).
template <typename T, int W> struct VecOf;
template <> struct VecOf<float, 4> { typedef float4 type; };
template <> struct VecOf<double, 2> { typedef double2 type; };

template <typename T, int W>
__device__ inline void vload(const T* __restrict__ p, T out[W]) {
  typedef typename VecOf<T, W>::type V;
  V v = *reinterpret_cast<const V*>(p);
  const T* e = reinterpret_cast<const T*>(&v);
#pragma unroll
  for (int k = 0; k < W; k++) out[k] = e[k];
}
template <typename T, int W>
__device__ inline void vstore(T* __restrict__ p, const T in[W]) {
  typedef typename VecOf<T, W>::type V;
  V v;
  T* e = reinterpret_cast<T*>(&v);
#pragma unroll
  for (int k = 0; k < W; k++) e[k] = in[k];
  *reinterpret_cast<V*>(p) = v;
}

// Bulk free-space cells, W consecutive x-cells per thread with 16-byte
// plane accesses (float4 in the precision form, double2 in the plain form):
// a 64-lane wavefront covers 64*W cells of a row.  Needs nx % W == 0 and a
// plane stride that is a multiple of W.  Per group of W cells:
//   own ports        12 vector loads
//   y/z neighbours   8 vector loads (the facing ports of rows y-1, y+1 and
//                    planes z-1, z+1)
//   x neighbours     inside the group from the group itself, across groups
//                    from the adjacent lanes, loads only at wavefront edges
// Groups whose cells are all bulk cells store with 12 vector stores; groups
// holding special cells store their bulk cells one by one.
template <typename T, bool PLAIN, bool PULL, int W>
__global__ __launch_bounds__(256) void k_bulk_vec(Dev d, const T* __restrict__ Lin, T* __restrict__ Lout,
                                                 PullArgs a, BulkMap m) {
  int bx, byi, z;
  bulk_tile(m, bx, byi, z);
  const int lane = threadIdx.x;
  const int x0 = (bx * 64 + lane) * W;
  const int y = byi * m.by + threadIdx.y;
  const bool valid = x0 < d.nx && y < d.ny && z < d.nz;
  const long long P = d.P;
  const long long sxy = (long long)d.nx * d.ny;
  const long long c = valid ? x0 + (long long)d.nx * (y + (long long)d.ny * z) : 0;
  T own[12][W];
  if (valid) {
#pragma unroll
    for (int p = 0; p < 12; p++) vload<T, W>(Lin + p * P + c, own[p]);
  } else {
#pragma unroll
    for (int p = 0; p < 12; p++)
#pragma unroll
      for (int k = 0; k < W; k++) own[p][k] = (T)0;
  }
  // x neighbours across groups (all lanes take part in the moves)
  T xa = __shfl_up(own[Vy_xmax][W - 1], 1, 64), xb = __shfl_up(own[Vz_xmax][W - 1], 1, 64);
  T xc = __shfl_down(own[Vy_xmin][0], 1, 64), xd = __shfl_down(own[Vz_xmin][0], 1, 64);
  if (!valid) return;
  uint32_t bits = 0;
  {
    const uint32_t wbits = d.mask[c >> 5];
    bits = (wbits >> (c & 31)) & ((1u << W) - 1u);  // W divides 32 and c % W == 0
  }
  if (bits == ((1u << W) - 1u)) return;  // all special
  T ya[2][W], yb[2][W], za[2][W], zb[2][W];  // [0]: min-side neighbour, [1]: max-side
  if (PULL) {
    if (lane == 0 && x0 > 0) { xa = Lin[Vy_xmax * P + c - 1]; xb = Lin[Vz_xmax * P + c - 1]; }
    if (lane == 63 && x0 + W < d.nx) { xc = Lin[Vy_xmin * P + c + W]; xd = Lin[Vz_xmin * P + c + W]; }
    if (y > 0) { vload<T, W>(Lin + Vx_ymax * P + c - d.nx, ya[0]); vload<T, W>(Lin + Vz_ymax * P + c - d.nx, yb[0]); }
    if (y + 1 < d.ny) { vload<T, W>(Lin + Vx_ymin * P + c + d.nx, ya[1]); vload<T, W>(Lin + Vz_ymin * P + c + d.nx, yb[1]); }
    if (z > 0) { vload<T, W>(Lin + Vx_zmax * P + c - sxy, za[0]); vload<T, W>(Lin + Vy_zmax * P + c - sxy, zb[0]); }
    if (z + 1 < d.nz) { vload<T, W>(Lin + Vx_zmin * P + c + sxy, za[1]); vload<T, W>(Lin + Vy_zmin * P + c + sxy, zb[1]); }
  }
  T out[12][W];
  // arithmetic type: FP64 in the plain form (bit-faithful), FP32 in the
  // precision form
  typedef typename std::conditional<PLAIN, double, T>::type A;
  const A Rx0 = a.apply_r ? (A)d.R[0] : (A)1, Rx1 = a.apply_r ? (A)d.R[1] : (A)1;
  const A Ry0 = a.apply_r ? (A)d.R[2] : (A)1, Ry1 = a.apply_r ? (A)d.R[3] : (A)1;
  const A Rz0 = a.apply_r ? (A)d.R[4] : (A)1, Rz1 = a.apply_r ? (A)d.R[5] : (A)1;
#pragma unroll
  for (int k = 0; k < W; k++) {
    const int x = x0 + k;
    A v[12];
#pragma unroll
    for (int p = 0; p < 12; p++) v[p] = (A)own[p][k];
    if (PULL) {
      const A nxa = k == 0 ? (A)xa : (A)own[Vy_xmax][k > 0 ? k - 1 : 0];
      const A nxb = k == 0 ? (A)xb : (A)own[Vz_xmax][k > 0 ? k - 1 : 0];
      const A nxc = k == W - 1 ? (A)xc : (A)own[Vy_xmin][k < W - 1 ? k + 1 : 0];
      const A nxd = k == W - 1 ? (A)xd : (A)own[Vz_xmin][k < W - 1 ? k + 1 : 0];
      if (x == 0) { v[Vy_xmin] = v[Vy_xmin] * Rx0; v[Vz_xmin] = v[Vz_xmin] * Rx0; }
      else { v[Vy_xmin] = free_face<PLAIN>(v[Vy_xmin], nxa); v[Vz_xmin] = free_face<PLAIN>(v[Vz_xmin], nxb); }
      if (x == d.nx - 1) { v[Vy_xmax] = v[Vy_xmax] * Rx1; v[Vz_xmax] = v[Vz_xmax] * Rx1; }
      else { v[Vy_xmax] = free_face<PLAIN>(v[Vy_xmax], nxc); v[Vz_xmax] = free_face<PLAIN>(v[Vz_xmax], nxd); }
      if (y == 0) { v[Vx_ymin] = v[Vx_ymin] * Ry0; v[Vz_ymin] = v[Vz_ymin] * Ry0; }
      else { v[Vx_ymin] = free_face<PLAIN>(v[Vx_ymin], (A)ya[0][k]); v[Vz_ymin] = free_face<PLAIN>(v[Vz_ymin], (A)yb[0][k]); }
      if (y == d.ny - 1) { v[Vx_ymax] = v[Vx_ymax] * Ry1; v[Vz_ymax] = v[Vz_ymax] * Ry1; }
      else { v[Vx_ymax] = free_face<PLAIN>(v[Vx_ymax], (A)ya[1][k]); v[Vz_ymax] = free_face<PLAIN>(v[Vz_ymax], (A)yb[1][k]); }
      if (z == 0) { v[Vx_zmin] = v[Vx_zmin] * Rz0; v[Vy_zmin] = v[Vy_zmin] * Rz0; }
      else { v[Vx_zmin] = free_face<PLAIN>(v[Vx_zmin], (A)za[0][k]); v[Vy_zmin] = free_face<PLAIN>(v[Vy_zmin], (A)zb[0][k]); }
      if (z == d.nz - 1) { v[Vx_zmax] = v[Vx_zmax] * Rz1; v[Vy_zmax] = v[Vy_zmax] * Rz1; }
      else { v[Vx_zmax] = free_face<PLAIN>(v[Vx_zmax], (A)za[1][k]); v[Vy_zmax] = free_face<PLAIN>(v[Vy_zmax], (A)zb[1][k]); }
    }
    if (PLAIN) scatter_free_plain((double*)v); else scatter_free_f32((float*)v);
#pragma unroll
    for (int p = 0; p < 12; p++) out[p][k] = (T)v[p];
  }
  if (bits == 0) {
#pragma unroll
    for (int p = 0; p < 12; p++) vstore<T, W>(Lout + p * P + c, out[p]);
  } else {
#pragma unroll
    for (int k = 0; k < W; k++)
      if (!((bits >> k) & 1u))
#pragma unroll
        for (int p = 0; p < 12; p++) Lout[p * P + c + k] = out[p][k];
  }
}

// Bulk free-space cells, marching zc planes per block.  The zmin-face ports
// (Vx_zmax, Vy_zmax of plane z-1) are the previous iteration's own values
// and the zmax-face ports (Vx_zmin, Vy_zmin of plane z+1) are kept for the
// next iteration, so each plane's links are read once per block; x and y
// neighbour ports are loaded (rows of the same tile or of tiles that run
// alongside on the same XCD).  Special cells keep the register pipeline
// but are not written.
template <typename T, bool PLAIN, bool PULL>
__global__ __launch_bounds__(256) void k_bulk_z(Dev d, const T* __restrict__ Lin, T* __restrict__ Lout,
                                               PullArgs a, BulkMap m) {
  int bx, byi, chunk;
  bulk_tile(m, bx, byi, chunk);
  const int x = bx * 64 + threadIdx.x;
  const int y = byi * m.by + threadIdx.y;
  if (x >= d.nx || y >= d.ny) return;
  const int z0 = chunk * m.zc;
  const int z1 = min(z0 + m.zc, d.nz);
  const long long P = d.P;
  const long long sxy = (long long)d.nx * d.ny;
  const long long c0 = x + (long long)d.nx * y;
  T pza = 0, pzb = 0;  // Vx_zmax, Vy_zmax of plane z-1
  T nza = 0, nzb = 0;  // Vx_zmin, Vy_zmin of plane z (loaded by the previous iteration)
  if (PULL && z0 > 0) {
    pza = Lin[Vx_zmax * P + c0 + (z0 - 1) * sxy];
    pzb = Lin[Vy_zmax * P + c0 + (z0 - 1) * sxy];
  }
  nza = Lin[Vx_zmin * P + c0 + z0 * sxy];
  nzb = Lin[Vy_zmin * P + c0 + z0 * sxy];
  for (int z = z0; z < z1; z++) {
    const long long c = c0 + z * sxy;
    double o[12];
#pragma unroll
    for (int p = 0; p < 12; p++)
      if (p != Vx_zmin && p != Vy_zmin) o[p] = (double)Lin[p * P + c];
    o[Vx_zmin] = (double)nza;
    o[Vy_zmin] = (double)nzb;
    T qa = 0, qb = 0;
    if (z + 1 < d.nz) {
      qa = Lin[Vx_zmin * P + c + sxy];
      qb = Lin[Vy_zmin * P + c + sxy];
    }
    const bool bulk = !((d.mask[c >> 5] >> (c & 31)) & 1u);
    if (bulk) {
      double v[12];
      if (PULL) {
        double xa = 0, xb = 0, xc = 0, xd = 0, ya = 0, yb = 0, yc = 0, yd = 0;
        if (x > 0) { xa = Lin[Vy_xmax * P + c - 1]; xb = Lin[Vz_xmax * P + c - 1]; }
        if (x + 1 < d.nx) { xc = Lin[Vy_xmin * P + c + 1]; xd = Lin[Vz_xmin * P + c + 1]; }
        if (y > 0) { ya = Lin[Vx_ymax * P + c - d.nx]; yb = Lin[Vz_ymax * P + c - d.nx]; }
        if (y + 1 < d.ny) { yc = Lin[Vx_ymin * P + c + d.nx]; yd = Lin[Vz_ymin * P + c + d.nx]; }
        bulk_incident<PLAIN>(d, a, x, y, z, o, xa, xb, xc, xd, ya, yb, yc, yd, (double)pza, (double)pzb,
                             (double)qa, (double)qb, v);
      } else {
#pragma unroll
        for (int p = 0; p < 12; p++) v[p] = o[p];
      }
      if (PLAIN) scatter_free_plain(v); else scatter_free_fast(v);
      store_cell(Lout, P, c, v);
    }
    pza = (T)o[Vx_zmax];
    pzb = (T)o[Vy_zmax];
    nza = qa;
    nzb = qb;
  }
}

// PML cells: scatter_PML.F90 stages 1-13.
template <typename T, bool PLAIN>
__device__ inline void pml_node(const Dev& d, double v[12], int slot, int i, int32_t cw, const PullArgs& a,
                                const double* wprev, const double* st = nullptr);
template <bool PLAIN>
__device__ inline void pml_state_preload(const Dev& d, int i, double st[9]);

template <typename T, bool PLAIN, bool PULL>
__device__ inline void pml_update(const Dev& d, const T* __restrict__ Lin, T* __restrict__ Lout,
                                  const PullArgs& a, int slot, int i) {
  const int32_t cw = d.pml_cell[i];
  const long long c = cw & kCellMask;
  double v[12];
  if (PULL && !(cw & kFlagFd) && a.prev_mode == 1) {
    // the common case: both pulls in one round trip
    double st[9];
    pml_state_preload<PLAIN>(d, i, st);
    int x, y, z, sg;
    cell_coords(d, c, x, y, z, sg);
    double w[12];
    bulk_pull_pair<T, PLAIN>(d, Lin, a, c, x, y, z, sg, v, w);
    flush_small(v, a.small);
    flush_small(w, a.prev_small);
    pml_node<T, PLAIN>(d, v, slot, i, cw, a, w, st);
    store_cell(Lout, d.P, c, v);
    return;
  }
  if (PULL) {
    int x, y, z, sg;
    cell_coords(d, c, x, y, z, sg);
    const int fdi = (cw & kFlagFd) ? sorted_lookup(d.pml_fd_tab, d.n_pml_fd, i) : -1;
    pull_cell<T, PLAIN>(d, Lin, c, x, y, z, sg, fdi, a, v);
  } else {
#pragma unroll
    for (int p = 0; p < 12; p++) v[p] = ld(Lin, d.P, p, c);
  }
  flush_small(v, a.small);
  pml_node<T, PLAIN>(d, v, slot, i, cw, a, nullptr);
  store_cell(Lout, d.P, c, v);
}

// A state word is read once per step and not again until the next one:
// non-temporal loads keep the state streams out of the caches (measured 3%
// off the PML kernel).
__device__ inline double state_ld(const double* p) { return __builtin_nontemporal_load(p); }

// The previous step's incident voltages of PML cell c (six ordinary faces),
// after that step's flush: what its pull produced then, recomputed from the
// link buffer it pulled from (a.prev), which no kernel has written since.
template <typename T, bool PLAIN>
__device__ inline void pml_prev_incident(const Dev& d, const PullArgs& a, long long c, double w[12]) {
  if (a.prev_mode == 0) {
#pragma unroll
    for (int p = 0; p < 12; p++) w[p] = 0.0;
    return;
  }
  const T* __restrict__ L = (const T*)a.prev;
  if (a.prev_mode == 1) {
    int x, y, z, sg;
    cell_coords(d, c, x, y, z, sg);
    PullArgs b{a.prev_apply_r, 0, 0.0, -1};
    bulk_pull<T, PLAIN>(d, L, b, c, x, y, z, sg, w);
  } else {
#pragma unroll
    for (int p = 0; p < 12; p++) w[p] = (double)lk_get(L, d.P, p, c);
  }
  flush_small(w, a.prev_small);
}

// scatter_PML.F90:22-313 stages 1-13 on the incident voltages v of PML list
// entry i (cw: its packed cell word), in place.  The previous step's scaled
// incident voltages (last_Vi) and branch currents (last_I) are recomputed
// from that step's incident voltages (pml_prev_incident) by the same
// operations that produced them, so they equal the stored words bit for bit;
// a cell with a special face keeps them as state words.  wprev: those
// incident voltages when the caller pulled them already (flushed), else null.
// The nine state words every PML cell carries (Vxt, Vyt, Vzt, Vyxt, Vzxt,
// Vxyt, Vzyt, Vyzt, Vxzt), loaded ahead so that they are in flight with the
// link loads.
// A 48-bit state word (reduced forms): the top 32 bits and the next 16 of
// the double's bit pattern; the low 16 bits read back as zero.  Word w of
// list position i sits at plane w (a uniform base) plus i: the per-lane
// part is a 32-bit byte offset shared by all nine words, so each access is
// a scalar base plus one offset register rather than a 64-bit address.
template <typename U>
__device__ inline U* s48_at(U* plane, uint32_t byte_off) {
  return (U*)((char*)plane + byte_off);
}
__device__ inline double s48_get(const Dev& d, int w, int i) {
  const long long n = d.n_pml_state;
  const uint64_t hi = __builtin_nontemporal_load(s48_at(d.pml_hi + w * n, (uint32_t)i * 4u));
  const uint64_t lo = __builtin_nontemporal_load(s48_at(d.pml_lo + w * n, (uint32_t)i * 2u));
  return __longlong_as_double((long long)((hi << 32) | (lo << 16)));
}
// Rounded to nearest on the magnitude (half away from zero): adding half a
// unit of the dropped 16 bits carries into the kept bits and, at the top of
// a binade, into the exponent, which is the correctly rounded value.
__device__ inline void s48_put(const Dev& d, int w, int i, double x) {
  const long long n = d.n_pml_state;
  const uint64_t b = (uint64_t)__double_as_longlong(x) + 0x8000ull;
  *s48_at(d.pml_hi + w * n, (uint32_t)i * 4u) = (uint32_t)(b >> 32);
  *s48_at(d.pml_lo + w * n, (uint32_t)i * 2u) = (uint16_t)(b >> 16);
}

template <bool PLAIN>
__device__ inline void pml_state_preload(const Dev& d, int i, double st[9]) {
  const long long n = d.n_pml_state;
  if (PLAIN) {
    const double* S = d.pml_state;
    const int w[9] = {PS_Vxt, PS_Vyt, PS_Vzt, PS_Vyxt, PS_Vzxt, PS_Vxyt, PS_Vzyt, PS_Vyzt, PS_Vxzt};
#pragma unroll
    for (int k = 0; k < 9; k++) st[k] = state_ld(&S[w[k] * n + i]);
  } else {
    // st[k] is word SW k
#pragma unroll
    for (int k = 0; k < 9; k++) st[k] = s48_get(d, k, i);
  }
}

template <typename T, bool PLAIN>
__device__ inline void pml_node(const Dev& d, double v[12], int slot, int i, int32_t cw, const PullArgs& a,
                                const double* wprev, const double* st) {
  const double Z0 = kZ0, dl = d.dl;
  const tlm_pml_parameter& Pp = d.pml_params[d.pml_par[i]];
  const double ax = Pp.ax, ay = Pp.ay, az = Pp.az;
  const double exp_x = Pp.exp_x, exp_y = Pp.exp_y, exp_z = Pp.exp_z;
  const double Csx = Pp.Csx, Csy = Pp.Csy, Csz = Pp.Csz;
  const double Csx2 = Pp.Csx2, Csy2 = Pp.Csy2, Csz2 = Pp.Csz2;
  const double Csy_sz = Pp.Csy_sz, Csz_sx = Pp.Csz_sx, Csx_sy = Pp.Csx_sy;
  const long long n = d.n_pml_state;
  const long long si = i;  // state slot = list position
  double* S = d.pml_state;
  // PML port aliases (PML_module.F90:60-71), 0-based
  enum { Vxny = 0, Vxnz, Vxpy, Vxpz, Vynx, Vynz, Vypx, Vypz, Vznx, Vzny, Vzpx, Vzpy };
  // The plain form keeps the reference's operations: I = curl / (2 Z0) and
  // the products with Z0 in their source order.  The reduced forms (not
  // bitwise) carry h = Z0 I = curl / 2 in place of every current, which
  // folds the division by 2 Z0 and the multiplications by Z0 away.
  auto cur = [&](double curl) { return PLAIN ? curl / (2.0 * Z0) : curl * 0.5; };
  auto zl = [&](double x) { return PLAIN ? Z0 * x : x; };   // Z0 * (x)
  auto zr = [&](double x) { return PLAIN ? x * Z0 : x; };   // (x) * Z0
  auto zm = [&](double a) { return PLAIN ? a * Z0 : a; };   // (a * Z0) * I
  // STAGE 2 (state reads)
  constexpr int wIx = PLAIN ? PS_Ix : PC_Ix, wIy = PLAIN ? PS_Iy : PC_Iy, wIz = PLAIN ? PS_Iz : PC_Iz;
  // previous incident voltages: recomputed (rec) or state words
  const bool rec = (cw & kFlagFd) == 0;
  double last_Ix, last_Iy, last_Iz;
  double lv[12];                               // plain: last_Vi
  double Szx, Syx, Sxy, Szy, Syz, Sxz;         // precision: pair sums of last_Vi
  if (rec) {
    double w[12];
    if (wprev) {
#pragma unroll
      for (int k = 0; k < 12; k++) w[k] = wprev[k];
    } else {
      pml_prev_incident<T, PLAIN>(d, a, cw & kCellMask, w);
    }
    w[Vynx] = w[Vynx] * exp_y; w[Vypx] = w[Vypx] * exp_y;
    w[Vznx] = w[Vznx] * exp_z; w[Vzpx] = w[Vzpx] * exp_z;
    w[Vxny] = w[Vxny] * exp_x; w[Vxpy] = w[Vxpy] * exp_x;
    w[Vzny] = w[Vzny] * exp_z; w[Vzpy] = w[Vzpy] * exp_z;
    w[Vxnz] = w[Vxnz] * exp_x; w[Vxpz] = w[Vxpz] * exp_x;
    w[Vynz] = w[Vynz] * exp_y; w[Vypz] = w[Vypz] * exp_y;
    last_Ix = cur(w[Vypz] - w[Vzpy] - w[Vynz] + w[Vzny]);
    last_Iy = cur(w[Vzpx] - w[Vxpz] - w[Vznx] + w[Vxnz]);
    last_Iz = cur(w[Vxpy] - w[Vypx] - w[Vxny] + w[Vynx]);
#pragma unroll
    for (int k = 0; k < 12; k++) lv[k] = w[k];
    Szx = w[Vznx] + w[Vzpx]; Syx = w[Vynx] + w[Vypx];
    Sxy = w[Vxny] + w[Vxpy]; Szy = w[Vzny] + w[Vzpy];
    Syz = w[Vynz] + w[Vypz]; Sxz = w[Vxnz] + w[Vxpz];
  } else {
    last_Ix = state_ld(&S[wIx * n + si]); last_Iy = state_ld(&S[wIy * n + si]); last_Iz = state_ld(&S[wIz * n + si]);
    if (PLAIN) {
#pragma unroll
      for (int k = 0; k < 12; k++) lv[k] = state_ld(&S[(PS_Vi + k) * n + si]);
      Szx = Syx = Sxy = Szy = Syz = Sxz = 0.0;
    } else {
#pragma unroll
      for (int k = 0; k < 12; k++) lv[k] = 0.0;
      Szx = state_ld(&S[PC_Szx * n + si]); Syx = state_ld(&S[PC_Syx * n + si]);
      Sxy = state_ld(&S[PC_Sxy * n + si]); Szy = state_ld(&S[PC_Szy * n + si]);
      Syz = state_ld(&S[PC_Syz * n + si]); Sxz = state_ld(&S[PC_Sxz * n + si]);
    }
  }
  // (loading these ahead of the previous step's pull was measured slower)
  double st_[9];
  if (!st) {
    pml_state_preload<PLAIN>(d, i, st_);
    st = st_;
  }
  const double last_Vxt = st[0], last_Vyt = st[1], last_Vzt = st[2];
  const double last_Vyxt = st[3], last_Vzxt = st[4];
  const double last_Vxyt = st[5], last_Vzyt = st[6];
  const double last_Vyzt = st[7], last_Vxzt = st[8];
  // STAGE 3
  v[Vynx] = v[Vynx] * exp_y; v[Vypx] = v[Vypx] * exp_y;
  v[Vznx] = v[Vznx] * exp_z; v[Vzpx] = v[Vzpx] * exp_z;
  v[Vxny] = v[Vxny] * exp_x; v[Vxpy] = v[Vxpy] * exp_x;
  v[Vzny] = v[Vzny] * exp_z; v[Vzpy] = v[Vzpy] * exp_z;
  v[Vxnz] = v[Vxnz] * exp_x; v[Vxpz] = v[Vxpz] * exp_x;
  v[Vynz] = v[Vynz] * exp_y; v[Vypz] = v[Vypz] * exp_y;
  // STAGE 4
  const double Vx = (v[Vynx] + v[Vypx] + v[Vznx] + v[Vzpx]) / 2.0;
  const double Vy = (v[Vxny] + v[Vxpy] + v[Vzny] + v[Vzpy]) / 2.0;
  const double Vz = (v[Vxnz] + v[Vxpz] + v[Vynz] + v[Vypz]) / 2.0;
  const double Ix = cur(v[Vypz] - v[Vzpy] - v[Vynz] + v[Vzny]);
  const double Iy = cur(v[Vzpx] - v[Vxpz] - v[Vznx] + v[Vxnz]);
  const double Iz = cur(v[Vxpy] - v[Vypx] - v[Vxny] + v[Vynx]);
  // STAGE 5
  double Vx_sh, Vy_sh, Vz_sh;
  if (PLAIN) {
    Vx_sh = ax * (-(lv[Vznx] + lv[Vzpx] + lv[Vynx] + lv[Vypx]) / 2.0
                  + Csy * ((v[Vznx] + v[Vzpx]) + (lv[Vznx] + lv[Vzpx]))
                  + Csz * ((v[Vynx] + v[Vypx]) + (lv[Vynx] + lv[Vypx]))
                  + Csy_sz * last_Vxt);
    Vy_sh = ay * (-(lv[Vxny] + lv[Vxpy] + lv[Vzny] + lv[Vzpy]) / 2.0
                  + Csz * ((v[Vxny] + v[Vxpy]) + (lv[Vxny] + lv[Vxpy]))
                  + Csx * ((v[Vzny] + v[Vzpy]) + (lv[Vzny] + lv[Vzpy]))
                  + Csz_sx * last_Vyt);
    Vz_sh = az * (-(lv[Vynz] + lv[Vypz] + lv[Vxnz] + lv[Vxpz]) / 2.0
                  + Csx * ((v[Vynz] + v[Vypz]) + (lv[Vynz] + lv[Vypz]))
                  + Csy * ((v[Vxnz] + v[Vxpz]) + (lv[Vxnz] + lv[Vxpz]))
                  + Csx_sy * last_Vzt);
  } else {
    Vx_sh = ax * (-(Szx + Syx) * 0.5 + Csy * ((v[Vznx] + v[Vzpx]) + Szx) + Csz * ((v[Vynx] + v[Vypx]) + Syx)
                  + Csy_sz * last_Vxt);
    Vy_sh = ay * (-(Sxy + Szy) * 0.5 + Csz * ((v[Vxny] + v[Vxpy]) + Sxy) + Csx * ((v[Vzny] + v[Vzpy]) + Szy)
                  + Csz_sx * last_Vyt);
    Vz_sh = az * (-(Syz + Sxz) * 0.5 + Csx * ((v[Vynz] + v[Vypz]) + Syz) + Csy * ((v[Vxnz] + v[Vxpz]) + Sxz)
                  + Csx_sy * last_Vzt);
  }
  // STAGE 6
  const double Vxt = ax * Vx + Vx_sh;
  const double Vyt = ay * Vy + Vy_sh;
  const double Vzt = az * Vz + Vz_sh;
  // STAGE 7a
  const int orow = (cw & kFlagOut) ? sorted_lookup(d.pml_out_tab, d.n_pml_out, i) : -1;
  double Ixt = 0.0, Iyt = 0.0, Izt = 0.0;
  if (orow >= 0) {
    const double lIxt = d.pml_It[0 * n + si], lIyt = d.pml_It[1 * n + si], lIzt = d.pml_It[2 * n + si];
    // currents proper (the reduced forms carry Z0 * I)
    const double ix = PLAIN ? Ix : Ix / Z0, iy = PLAIN ? Iy : Iy / Z0, iz = PLAIN ? Iz : Iz / Z0;
    const double lix = PLAIN ? last_Ix : last_Ix / Z0, liy = PLAIN ? last_Iy : last_Iy / Z0;
    const double liz = PLAIN ? last_Iz : last_Iz / Z0;
    Ixt = ax * (ix - lix + Csx2 * (ix + lix) + Csy_sz * lIxt);
    Iyt = ay * (iy - liy + Csy2 * (iy + liy) + Csz_sx * lIyt);
    Izt = az * (iz - liz + Csz2 * (iz + liz) + Csx_sy * lIzt);
  }
  // STAGE 7
  const double Vxy_se = az * (zl(-last_Iz + Csy2 * (Iz + last_Iz)) + Csx_sy * last_Vxyt);
  const double Vyx_se = az * (zl(-last_Iz + Csx2 * (Iz + last_Iz)) + Csx_sy * last_Vyxt);
  const double Vzx_se = ay * (zl(-last_Iy + Csx2 * (Iy + last_Iy)) + Csz_sx * last_Vzxt);
  const double Vxz_se = ay * (zl(-last_Iy + Csz2 * (Iy + last_Iy)) + Csz_sx * last_Vxzt);
  const double Vyz_se = ax * (zl(-last_Ix + Csz2 * (Ix + last_Ix)) + Csy_sz * last_Vyzt);
  const double Vzy_se = ax * (zl(-last_Ix + Csy2 * (Ix + last_Ix)) + Csy_sz * last_Vzyt);
  // STAGE 8
  const double Vxyt = zr(az * Iz) + Vxy_se;
  const double Vyxt = zr(az * Iz) + Vyx_se;
  const double Vzxt = zr(ay * Iy) + Vzx_se;
  const double Vxzt = zr(ay * Iy) + Vxz_se;
  const double Vyzt = zr(ax * Ix) + Vyz_se;
  const double Vzyt = zr(ax * Ix) + Vzy_se;
  // STAGE 9
  if (!rec) {
    S[wIx * n + si] = Ix; S[wIy * n + si] = Iy; S[wIz * n + si] = Iz;
    if (PLAIN) {
#pragma unroll
      for (int k = 0; k < 12; k++) S[(PS_Vi + k) * n + si] = v[k];
    } else {
      S[PC_Szx * n + si] = v[Vznx] + v[Vzpx]; S[PC_Syx * n + si] = v[Vynx] + v[Vypx];
      S[PC_Sxy * n + si] = v[Vxny] + v[Vxpy]; S[PC_Szy * n + si] = v[Vzny] + v[Vzpy];
      S[PC_Syz * n + si] = v[Vynz] + v[Vypz]; S[PC_Sxz * n + si] = v[Vxnz] + v[Vxpz];
    }
  }
  if (PLAIN) {
    S[PS_Vxt * n + si] = Vxt; S[PS_Vyt * n + si] = Vyt; S[PS_Vzt * n + si] = Vzt;
  } else {
    s48_put(d, SW_Vxt, i, Vxt); s48_put(d, SW_Vyt, i, Vyt); s48_put(d, SW_Vzt, i, Vzt);
  }
  if (orow >= 0) {
    d.pml_It[0 * n + si] = Ixt; d.pml_It[1 * n + si] = Iyt; d.pml_It[2 * n + si] = Izt;
  }
  if (PLAIN) {
    S[PS_Vxyt * n + si] = Vxyt; S[PS_Vyxt * n + si] = Vyxt;
    S[PS_Vzxt * n + si] = Vzxt; S[PS_Vxzt * n + si] = Vxzt;
    S[PS_Vzyt * n + si] = Vzyt; S[PS_Vyzt * n + si] = Vyzt;
  } else {
    s48_put(d, SW_Vxyt, i, Vxyt); s48_put(d, SW_Vyxt, i, Vyxt);
    s48_put(d, SW_Vzxt, i, Vzxt); s48_put(d, SW_Vxzt, i, Vxzt);
    s48_put(d, SW_Vzyt, i, Vzyt); s48_put(d, SW_Vyzt, i, Vyzt);
  }
  // STAGE 10
  const double Vynx_P = Vx_sh - Vyx_se, Vypx_P = Vx_sh + Vyx_se;
  const double Vznx_P = Vx_sh + Vzx_se, Vzpx_P = Vx_sh - Vzx_se;
  const double Vxny_P = Vy_sh + Vxy_se, Vxpy_P = Vy_sh - Vxy_se;
  const double Vzny_P = Vy_sh - Vzy_se, Vzpy_P = Vy_sh + Vzy_se;
  const double Vynz_P = Vz_sh + Vyz_se, Vypz_P = Vz_sh - Vyz_se;
  const double Vxnz_P = Vz_sh - Vxz_se, Vxpz_P = Vz_sh + Vxz_se;
  // STAGE 11
  double M, m;
  M = v[Vypx]; m = v[Vynx];
  v[Vynx] = ax * Vx - zm(az) * Iz - M + Vynx_P;
  v[Vypx] = ax * Vx + zm(az) * Iz - m + Vypx_P;
  M = v[Vzpx]; m = v[Vznx];
  v[Vznx] = ax * Vx + zm(ay) * Iy - M + Vznx_P;
  v[Vzpx] = ax * Vx - zm(ay) * Iy - m + Vzpx_P;
  M = v[Vzpy]; m = v[Vzny];
  v[Vzny] = ay * Vy - zm(ax) * Ix - M + Vzny_P;
  v[Vzpy] = ay * Vy + zm(ax) * Ix - m + Vzpy_P;
  M = v[Vxpy]; m = v[Vxny];
  v[Vxny] = ay * Vy + zm(az) * Iz - M + Vxny_P;
  v[Vxpy] = ay * Vy - zm(az) * Iz - m + Vxpy_P;
  M = v[Vxpz]; m = v[Vxnz];
  v[Vxnz] = az * Vz - zm(ay) * Iy - M + Vxnz_P;
  v[Vxpz] = az * Vz + zm(ay) * Iy - m + Vxpz_P;
  M = v[Vypz]; m = v[Vynz];
  v[Vynz] = az * Vz + zm(ax) * Ix - M + Vynz_P;
  v[Vypz] = az * Vz - zm(ax) * Ix - m + Vypz_P;
  // STAGE 12
  if (orow >= 0) {
    double* fo = d.ring + ((long long)slot * d.n_out_rows + orow) * 6;
    fo[0] = -Vxt / dl; fo[1] = -Vyt / dl; fo[2] = -Vzt / dl;
    fo[3] = Ixt / dl; fo[4] = Iyt / dl; fo[5] = Izt / dl;
  }
  // STAGE 13
  v[Vynx] = v[Vynx] * exp_y; v[Vypx] = v[Vypx] * exp_y;
  v[Vznx] = v[Vznx] * exp_z; v[Vzpx] = v[Vzpx] * exp_z;
  v[Vxny] = v[Vxny] * exp_x; v[Vxpy] = v[Vxpy] * exp_x;
  v[Vzny] = v[Vzny] * exp_z; v[Vzpy] = v[Vzpy] * exp_z;
  v[Vxnz] = v[Vxnz] * exp_x; v[Vxpz] = v[Vxpz] * exp_x;
  v[Vynz] = v[Vynz] * exp_y; v[Vypz] = v[Vypz] * exp_y;
}

// A PML cell without output row or special face, list position i, cell
// c0 + k of a chunk (c0 the same in every lane) at (x, y, z) of segment sg:
// pml_update's operations with the chunk's addressing.
template <typename T, bool PLAIN, bool PULL>
__device__ inline void pml_fast_cell(const Dev& d, const T* __restrict__ Lin, T* __restrict__ Lout,
                                     const PullArgs& a, int slot, int i, long long c0, uint32_t k, int x, int y,
                                     int z, int sg) {
  double st[9];
  pml_state_preload<PLAIN>(d, i, st);
  double v[12], wv[12];
  if (PULL && a.prev_mode == 1) {
    // the common step: both pulls in one round trip
    chunk_pull_pair<T, PLAIN>(d, Lin, a, c0, k, x, y, z, sg, v, wv);
  } else {
    // the first steps: this step's incident voltages pulled (or, without a
    // pending connect, the buffer's own values), and the previous step's as
    // pml_prev_incident gives them (none, pulled, or the buffer's own values)
    if (PULL) {
      chunk_pull<T, PLAIN>(d, Lin, a, c0, k, x, y, z, sg, v);
    } else {
#pragma unroll
      for (int p = 0; p < 12; p++) v[p] = (double)lk_get_u(Lin, d.P, p, c0, k);
    }
    const T* __restrict__ Q = (const T*)a.prev;
    if (a.prev_mode == 1) {
      PullArgs b{a.prev_apply_r, 0, 0.0, -1};
      chunk_pull<T, PLAIN>(d, Q, b, c0, k, x, y, z, sg, wv);
    } else if (a.prev_mode == 2) {
#pragma unroll
      for (int p = 0; p < 12; p++) wv[p] = (double)lk_get_u(Q, d.P, p, c0, k);
    } else {
#pragma unroll
      for (int p = 0; p < 12; p++) wv[p] = 0.0;
    }
  }
  flush_small(v, a.small);
  flush_small(wv, a.prev_small);
  // a cell word without flags: no output row, no special face
  pml_node<T, PLAIN>(d, v, slot, i, (int32_t)(c0 + k) & kCellMask, a, wv, st);
  store_links_u(Lout, d.P, c0, k, v);
}

// PML cells of pure PML lines from per-chunk records (two int4 each, read
// by all lanes at one address):
//   e = (chunk << 4 | mask of the chunk's pure PML lines,
//        x of the chunk's first cell | segment << 16, y, z)
//   f = (list position of the first of those cells, 1 if any of them has an
//        output row or a special face, 0, 0)
// The pure-line cells of the PML list are in storage order, so a lane's
// list position follows from the lines of the mask below its own.  Records
// come in two lists: chunks without flagged cells, and chunks with one.
// FAST = true runs the first list: cell and coordinates from the record, no
// list word, no division for a segment at least 64 cells wide, and no
// special-face or output path compiled in (126 against 149 registers: four
// waves per SIMD instead of three).  Its incident voltages come as
// pml_update's do in every step kind (pending connect or not; previous
// buffer pulled, read as it is, or absent).  FAST = false runs pml_update
// on each lane's list position (the flagged chunks).  Padding
// lanes (after a segment's last row, in the last line of a segment) have no
// list position: they leave as soon as their coordinates show it.
template <typename T, bool PLAIN, bool PULL, bool FAST>
__device__ inline void pml_chunk_cell(const Dev& d, const T* __restrict__ Lin, T* __restrict__ Lout,
                                      const PullArgs& a, int slot, int4 e, int4 f, int lane) {
  const int L = d.line_len;
  const int q = lane / L;
  if (!((e.x >> q) & 1)) return;
  const int i = f.x + L * __popc((unsigned)(e.x & ((1 << q) - 1))) + (lane - q * L);
  const long long c = (long long)(e.x >> 4) * kChunk + lane;
  int sg = e.y >> 16, x = (e.y & 0xffff) + lane, y = e.z, z = e.w;
  const int w = seg_width(d, sg);
  if (w >= kChunk) {
    if (x >= d.seg_x0[sg + 1]) {
      x -= w;
      if (++y == d.ny) { y = 0; z++; }
    }
    if (z >= d.nz) return;  // padding after the segment's last row
  } else {
    // narrow segment (a PML slab a few cells thick): rows of w cells
    const unsigned r = (unsigned)(x - d.seg_x0[sg]), row = r / (unsigned)w;
    x = d.seg_x0[sg] + (int)(r - row * (unsigned)w);
    const unsigned yy = (unsigned)y + row, dz = yy / (unsigned)d.ny;
    y = (int)(yy - dz * (unsigned)d.ny);
    z += (int)dz;
    if (z >= d.nz) return;
  }
  if (!FAST) {
    pml_update<T, PLAIN, PULL>(d, Lin, Lout, a, slot, i);
    return;
  }
  pml_fast_cell<T, PLAIN, PULL>(d, Lin, Lout, a, slot, i, (long long)(e.x >> 4) * kChunk, (uint32_t)lane, x, y, z,
                                sg);
}

template <typename T, bool PLAIN, bool PULL, bool FAST>
__global__ __launch_bounds__(kChunk) void k_pml_chunks(Dev d, const T* __restrict__ Lin, T* __restrict__ Lout,
                                                       PullArgs a, int slot, const int4* __restrict__ rec) {
  pml_chunk_cell<T, PLAIN, PULL, FAST>(d, Lin, Lout, a, slot, rec[2 * blockIdx.x], rec[2 * blockIdx.x + 1],
                                       threadIdx.x);
}


// PML cell holding a non-dispersive material, scatter_PML_material.F90:22-534
// (two-step recursions for the stretched node voltages and currents, a
// capacitive and an inductive stub).  Same operation order as the
// reference; the STAGE 1 coefficients are recomputed per cell as there.
template <typename T, bool PLAIN, bool PULL>
__device__ inline void pml_material_update(const Dev& d, const T* __restrict__ Lin, T* __restrict__ Lout,
                                           const PullArgs& a, int slot, int i) {
  const long long c = d.pm_cell[i];
  double v[12];
  if (PULL) {
    int x, y, z, sg;
    cell_coords(d, c, x, y, z, sg);
    pull_cell<T, PLAIN>(d, Lin, c, x, y, z, sg, d.pm_fd[i], a, v);
  } else {
#pragma unroll
    for (int p = 0; p < 12; p++) v[p] = ld(Lin, d.P, p, c);
  }
  flush_small(v, a.small);
  enum { Vxny = 0, Vxnz, Vxpy, Vxpz, Vynx, Vynz, Vypx, Vypz, Vznx, Vzny, Vzpx, Vzpy };
  const double Z0 = kZ0, dl = d.dl, dt = d.dt;
  const tlm_volume_material& vm = d.vmat[d.pm_mat[i] - 1];
  const tlm_pml_parameter& Pp = d.pml_params[d.pm_par[i]];
  // STAGE 1
  const double Ysx = vm.Ys_eps_f / kY0, Ysy = vm.Ys_eps_f / kY0, Ysz = vm.Ys_eps_f / kY0;
  const double Gex = vm.Ge / kY0, Gey = vm.Ge / kY0, Gez = vm.Ge / kY0;
  const double Zsx = vm.Zs_mu_f, Zsy = vm.Zs_mu_f, Zsz = vm.Zs_mu_f;
  const double Rmx = vm.Rm, Rmy = vm.Rm, Rmz = vm.Rm;
  const double sx = Pp.sx, sy = Pp.sy, sz = Pp.sz;
  const double exp_x = Pp.exp_x, exp_y = Pp.exp_y, exp_z = Pp.exp_z;
  double GpY2;
  GpY2 = (Gex + Ysx) / 2.0;
  const double Ax0 = 8.0 + 4.0 * GpY2 + 2.0 * dt * (sy + sz) * (1.0 + GpY2) + dt * dt * sy * sz * GpY2;
  const double Ax1 = -2.0 * (8.0 + 4.0 * GpY2) + 2.0 * (dt * dt * sy * sz * GpY2);
  const double Ax2 = 8.0 + 4.0 * GpY2 - 2.0 * dt * (sy + sz) * (1.0 + GpY2) + dt * dt * sy * sz * GpY2;
  GpY2 = (Gey + Ysy) / 2.0;
  const double Ay0 = 8.0 + 4.0 * GpY2 + 2.0 * dt * (sz + sx) * (1.0 + GpY2) + dt * dt * sz * sx * GpY2;
  const double Ay1 = -2.0 * (8.0 + 4.0 * GpY2) + 2.0 * (dt * dt * sz * sx * GpY2);
  const double Ay2 = 8.0 + 4.0 * GpY2 - 2.0 * dt * (sz + sx) * (1.0 + GpY2) + dt * dt * sz * sx * GpY2;
  GpY2 = (Gez + Ysz) / 2.0;
  const double Az0 = 8.0 + 4.0 * GpY2 + 2.0 * dt * (sx + sy) * (1.0 + GpY2) + dt * dt * sx * sy * GpY2;
  const double Az1 = -2.0 * (8.0 + 4.0 * GpY2) + 2.0 * (dt * dt * sx * sy * GpY2);
  const double Az2 = 8.0 + 4.0 * GpY2 - 2.0 * dt * (sx + sy) * (1.0 + GpY2) + dt * dt * sx * sy * GpY2;
  const double Bx0 = 4.0 + 2.0 * dt * sy, Bx1 = -8.0, Bx2 = 4.0 - 2.0 * dt * sy;
  const double By0 = 4.0 + 2.0 * dt * sz, By1 = -8.0, By2 = 4.0 - 2.0 * dt * sz;
  const double Bz0 = 4.0 + 2.0 * dt * sx, Bz1 = -8.0, Bz2 = 4.0 - 2.0 * dt * sx;
  const double Cx0 = 4.0 + 2.0 * dt * sz, Cx1 = -8.0, Cx2 = 4.0 - 2.0 * dt * sz;
  const double Cy0 = 4.0 + 2.0 * dt * sx, Cy1 = -8.0, Cy2 = 4.0 - 2.0 * dt * sx;
  const double Cz0 = 4.0 + 2.0 * dt * sy, Cz1 = -8.0, Cz2 = 4.0 - 2.0 * dt * sy;
  const double Dx0 = 4.0 * Ysx + 2.0 * dt * (sy + sz) * Ysx + 1.0 * (dt * dt * sy * sz * Ysx);
  const double Dx1 = -8.0 * Ysx + 2.0 * (dt * dt * sy * sz * Ysx);
  const double Dx2 = 4.0 * Ysx - 2.0 * dt * (sy + sz) * Ysx + 1.0 * (dt * dt * sy * sz * Ysx);
  const double Dy0 = 4.0 * Ysy + 2.0 * dt * (sz + sx) * Ysy + 1.0 * (dt * dt * sz * sx * Ysy);
  const double Dy1 = -8.0 * Ysy + 2.0 * (dt * dt * sz * sx * Ysy);
  const double Dy2 = 4.0 * Ysy - 2.0 * dt * (sz + sx) * Ysy + 1.0 * (dt * dt * sz * sx * Ysy);
  const double Dz0 = 4.0 * Ysz + 2.0 * dt * (sx + sy) * Ysz + 1.0 * (dt * dt * sx * sy * Ysz);
  const double Dz1 = -8.0 * Ysz + 2.0 * (dt * dt * sx * sy * Ysz);
  const double Dz2 = 4.0 * Ysz - 2.0 * dt * (sx + sy) * Ysz + 1.0 * (dt * dt * sx * sy * Ysz);
  double ZpR;
  ZpR = Zsx + Rmx;
  const double Sx0 = 16.0 * Z0 + 4.0 * ZpR + 2.0 * dt * (sy + sz) * (2.0 * Z0 + ZpR) + dt * dt * sy * sz * ZpR;
  const double Sx1 = -32.0 * Z0 - 8.0 * ZpR + 2.0 * dt * dt * sy * sz * ZpR;
  const double Sx2 = 16.0 * Z0 + 4.0 * ZpR - 2.0 * dt * (sy + sz) * (2.0 * Z0 + ZpR) + dt * dt * sy * sz * ZpR;
  ZpR = Zsy + Rmy;
  const double Sy0 = 16.0 * Z0 + 4.0 * ZpR + 2.0 * dt * (sz + sx) * (2.0 * Z0 + ZpR) + dt * dt * sz * sx * ZpR;
  const double Sy1 = -32.0 * Z0 - 8.0 * ZpR + 2.0 * dt * dt * sz * sx * ZpR;
  const double Sy2 = 16.0 * Z0 + 4.0 * ZpR - 2.0 * dt * (sz + sx) * (2.0 * Z0 + ZpR) + dt * dt * sz * sx * ZpR;
  ZpR = Zsz + Rmz;
  const double Sz0 = 16.0 * Z0 + 4.0 * ZpR + 2.0 * dt * (sx + sy) * (2.0 * Z0 + ZpR) + dt * dt * sx * sy * ZpR;
  const double Sz1 = -32.0 * Z0 - 8.0 * ZpR + 2.0 * dt * dt * sx * sy * ZpR;
  const double Sz2 = 16.0 * Z0 + 4.0 * ZpR - 2.0 * dt * (sx + sy) * (2.0 * Z0 + ZpR) + dt * dt * sx * sy * ZpR;
  const double Tx0 = 8.0 + 4.0 * dt * sx, Tx1 = -16.0, Tx2 = 8.0 - 4.0 * dt * sx;
  const double Ty0 = 8.0 + 4.0 * dt * sy, Ty1 = -16.0, Ty2 = 8.0 - 4.0 * dt * sy;
  const double Tz0 = 8.0 + 4.0 * dt * sz, Tz1 = -16.0, Tz2 = 8.0 - 4.0 * dt * sz;
  const double Uxy0 = 2.0 + dt * sz, Uxy1 = -2.0 + dt * sz, Wxy0 = Z0 * (2.0 + dt * sy), Wxy1 = Z0 * (-2.0 + dt * sy);
  const double Uyx0 = 2.0 + dt * sz, Uyx1 = -2.0 + dt * sz, Wyx0 = Z0 * (2.0 + dt * sx), Wyx1 = Z0 * (-2.0 + dt * sx);
  const double Uyz0 = 2.0 + dt * sx, Uyz1 = -2.0 + dt * sx, Wyz0 = Z0 * (2.0 + dt * sz), Wyz1 = Z0 * (-2.0 + dt * sz);
  const double Uzy0 = 2.0 + dt * sx, Uzy1 = -2.0 + dt * sx, Wzy0 = Z0 * (2.0 + dt * sy), Wzy1 = Z0 * (-2.0 + dt * sy);
  const double Uzx0 = 2.0 + dt * sy, Uzx1 = -2.0 + dt * sy, Wzx0 = Z0 * (2.0 + dt * sx), Wzx1 = Z0 * (-2.0 + dt * sx);
  const double Uxz0 = 2.0 + dt * sy, Uxz1 = -2.0 + dt * sy, Wxz0 = Z0 * (2.0 + dt * sz), Wxz1 = Z0 * (-2.0 + dt * sz);
  const double Px0 = 4.0 + 2.0 * dt * sx, Px1 = -8.0, Px2 = 4.0 - 2.0 * dt * sx;
  const double Py0 = 4.0 + 2.0 * dt * sy, Py1 = -8.0, Py2 = 4.0 - 2.0 * dt * sy;
  const double Pz0 = 4.0 + 2.0 * dt * sz, Pz1 = -8.0, Pz2 = 4.0 - 2.0 * dt * sz;
  const double Qx0 = 4.0 * Zsx + 2.0 * dt * (sy + sz) * Zsx + 1.0 * (dt * dt * sy * sz * Zsx);
  const double Qx1 = -8.0 * Zsx + 2.0 * (dt * dt * sy * sz * Zsx);
  const double Qx2 = 4.0 * Zsx - 2.0 * dt * (sy + sz) * Zsx + 1.0 * (dt * dt * sy * sz * Zsx);
  const double Qy0 = 4.0 * Zsy + 2.0 * dt * (sz + sx) * Zsy + 1.0 * (dt * dt * sz * sx * Zsy);
  const double Qy1 = -8.0 * Zsy + 2.0 * (dt * dt * sz * sx * Zsy);
  const double Qy2 = 4.0 * Zsy - 2.0 * dt * (sz + sx) * Zsy + 1.0 * (dt * dt * sz * sx * Zsy);
  const double Qz0 = 4.0 * Zsz + 2.0 * dt * (sx + sy) * Zsz + 1.0 * (dt * dt * sx * sy * Zsz);
  const double Qz1 = -8.0 * Zsz + 2.0 * (dt * dt * sx * sy * Zsz);
  const double Qz2 = 4.0 * Zsz - 2.0 * dt * (sx + sy) * Zsz + 1.0 * (dt * dt * sx * sy * Zsz);
  // STAGE 2
  const long long n = d.n_pm;
  double* M = d.pm_state;
  auto W = [&](int w) -> double& { return M[(long long)w * n + i]; };
  const double Ixt_m1 = W(PM_Ixt), Iyt_m1 = W(PM_Iyt), Izt_m1 = W(PM_Izt);
  const double Vxt_m1 = W(PM_Vxt), Vyt_m1 = W(PM_Vyt), Vzt_m1 = W(PM_Vzt);
  const double Vyxt_m1 = W(PM_Vyxt), Vzxt_m1 = W(PM_Vzxt), Vxyt_m1 = W(PM_Vxyt);
  const double Vzyt_m1 = W(PM_Vzyt), Vyzt_m1 = W(PM_Vyzt), Vxzt_m1 = W(PM_Vxzt);
  const double Ixt_m2 = W(PM_Ixt_m1), Iyt_m2 = W(PM_Iyt_m1), Izt_m2 = W(PM_Izt_m1);
  const double Vcx_m2 = W(PM_Vcx_m1), Vcy_m2 = W(PM_Vcy_m1), Vcz_m2 = W(PM_Vcz_m1);
  const double Vcx_m1 = W(PM_Vcx), Vcy_m1 = W(PM_Vcy), Vcz_m1 = W(PM_Vcz);
  const double Vxt_m2 = W(PM_Vxt_m1), Vyt_m2 = W(PM_Vyt_m1), Vzt_m2 = W(PM_Vzt_m1);
  const double Vosx = W(PM_Vosxr), Vosy = W(PM_Vosyr), Vosz = W(PM_Voszr);
  const double Vosx_m1 = W(PM_Vosx), Vosx_m2 = W(PM_Vosx_m1);
  const double Vosy_m1 = W(PM_Vosy), Vosy_m2 = W(PM_Vosy_m1);
  const double Vosz_m1 = W(PM_Vosz), Vosz_m2 = W(PM_Vosz_m1);
  const double Vssx = -W(PM_Vssxr), Vssy = -W(PM_Vssyr), Vssz = -W(PM_Vsszr);
  const double Vsx_m1 = W(PM_Vsx), Vsx_m2 = W(PM_Vsx_m1);
  const double Vsy_m1 = W(PM_Vsy), Vsy_m2 = W(PM_Vsy_m1);
  const double Vsz_m1 = W(PM_Vsz), Vsz_m2 = W(PM_Vsz_m1);
  auto Vi1 = [&](int p) { return W(PM_Vi + p); };
  auto Vi2 = [&](int p) { return W(PM_Vi_m1 + p); };
  // STAGE 3
  v[Vynx] = v[Vynx] * exp_y; v[Vypx] = v[Vypx] * exp_y;
  v[Vznx] = v[Vznx] * exp_z; v[Vzpx] = v[Vzpx] * exp_z;
  v[Vxny] = v[Vxny] * exp_x; v[Vxpy] = v[Vxpy] * exp_x;
  v[Vzny] = v[Vzny] * exp_z; v[Vzpy] = v[Vzpy] * exp_z;
  v[Vxnz] = v[Vxnz] * exp_x; v[Vxpz] = v[Vxpz] * exp_x;
  v[Vynz] = v[Vynz] * exp_y; v[Vypz] = v[Vypz] * exp_y;
  // STAGE 4
  const double Vxt = (1.0 / Ax0) * (Bx0 * (v[Vznx] + v[Vzpx]) + Cx0 * (v[Vynx] + v[Vypx]) + Dx0 * (Vosx)
                                    + Bx1 * (Vi1(Vznx) + Vi1(Vzpx)) + Cx1 * (Vi1(Vynx) + Vi1(Vypx)) + Dx1 * (Vosx_m1)
                                    + Bx2 * (Vi2(Vznx) + Vi2(Vzpx)) + Cx2 * (Vi2(Vynx) + Vi2(Vypx)) + Dx2 * (Vosx_m2)
                                    - Ax1 * Vxt_m1 - Ax2 * Vxt_m2);
  const double Vyt = (1.0 / Ay0) * (By0 * (v[Vxny] + v[Vxpy]) + Cy0 * (v[Vzny] + v[Vzpy]) + Dy0 * (Vosy)
                                    + By1 * (Vi1(Vxny) + Vi1(Vxpy)) + Cy1 * (Vi1(Vzny) + Vi1(Vzpy)) + Dy1 * (Vosy_m1)
                                    + By2 * (Vi2(Vxny) + Vi2(Vxpy)) + Cy2 * (Vi2(Vzny) + Vi2(Vzpy)) + Dy2 * (Vosy_m2)
                                    - Ay1 * Vyt_m1 - Ay2 * Vyt_m2);
  const double Vzt = (1.0 / Az0) * (Bz0 * (v[Vynz] + v[Vypz]) + Cz0 * (v[Vxnz] + v[Vxpz]) + Dz0 * (Vosz)
                                    + Bz1 * (Vi1(Vynz) + Vi1(Vypz)) + Cz1 * (Vi1(Vxnz) + Vi1(Vxpz)) + Dz1 * (Vosz_m1)
                                    + Bz2 * (Vi2(Vynz) + Vi2(Vypz)) + Cz2 * (Vi2(Vxnz) + Vi2(Vxpz)) + Dz2 * (Vosz_m2)
                                    - Az1 * Vzt_m1 - Az2 * Vzt_m2);
  // STAGE 5
  const double Vcx = v[Vypz] - v[Vzpy] - v[Vynz] + v[Vzny] - Vssx;
  const double Vcy = v[Vzpx] - v[Vxpz] - v[Vznx] + v[Vxnz] - Vssy;
  const double Vcz = v[Vxpy] - v[Vypx] - v[Vxny] + v[Vynx] - Vssz;
  const double Ixt = (1.0 / Sx0) * (Tx0 * Vcx + Tx1 * Vcx_m1 + Tx2 * Vcx_m2 - Sx1 * Ixt_m1 - Sx2 * Ixt_m2);
  const double Iyt = (1.0 / Sy0) * (Ty0 * Vcy + Ty1 * Vcy_m1 + Ty2 * Vcy_m2 - Sy1 * Iyt_m1 - Sy2 * Iyt_m2);
  const double Izt = (1.0 / Sz0) * (Tz0 * Vcz + Tz1 * Vcz_m1 + Tz2 * Vcz_m2 - Sz1 * Izt_m1 - Sz2 * Izt_m2);
  const double Vxyt = (1.0 / Uxy0) * (Wxy0 * Izt + Wxy1 * Izt_m1 - Uxy1 * Vxyt_m1);
  const double Vyxt = (1.0 / Uyx0) * (Wyx0 * Izt + Wyx1 * Izt_m1 - Uyx1 * Vyxt_m1);
  const double Vyzt = (1.0 / Uyz0) * (Wyz0 * Ixt + Wyz1 * Ixt_m1 - Uyz1 * Vyzt_m1);
  const double Vzyt = (1.0 / Uzy0) * (Wzy0 * Ixt + Wzy1 * Ixt_m1 - Uzy1 * Vzyt_m1);
  const double Vxzt = (1.0 / Uxz0) * (Wxz0 * Iyt + Wxz1 * Iyt_m1 - Uxz1 * Vxzt_m1);
  const double Vzxt = (1.0 / Uzx0) * (Wzx0 * Iyt + Wzx1 * Iyt_m1 - Uzx1 * Vzxt_m1);
  // STAGE 9: this step's values in, the previous step's to the _m1 slots
#pragma unroll
  for (int p = 0; p < 12; p++) {
    W(PM_Vi_m1 + p) = W(PM_Vi + p);
    W(PM_Vi + p) = v[p];
  }
  W(PM_Vxt_m1) = Vxt_m1; W(PM_Vyt_m1) = Vyt_m1; W(PM_Vzt_m1) = Vzt_m1;
  W(PM_Vxt) = Vxt; W(PM_Vyt) = Vyt; W(PM_Vzt) = Vzt;
  W(PM_Ixt_m1) = Ixt_m1; W(PM_Iyt_m1) = Iyt_m1; W(PM_Izt_m1) = Izt_m1;
  W(PM_Ixt) = Ixt; W(PM_Iyt) = Iyt; W(PM_Izt) = Izt;
  W(PM_Vcx_m1) = Vcx_m1; W(PM_Vcy_m1) = Vcy_m1; W(PM_Vcz_m1) = Vcz_m1;
  W(PM_Vcx) = Vcx; W(PM_Vcy) = Vcy; W(PM_Vcz) = Vcz;
  W(PM_Vxyt_m1) = Vxyt_m1; W(PM_Vyxt_m1) = Vyxt_m1;
  W(PM_Vxyt) = Vxyt; W(PM_Vyxt) = Vyxt;
  W(PM_Vzxt_m1) = Vzxt_m1; W(PM_Vxzt_m1) = Vxzt_m1;
  W(PM_Vzxt) = Vzxt; W(PM_Vxzt) = Vxzt;
  W(PM_Vzyt_m1) = Vzyt_m1; W(PM_Vyzt_m1) = Vyzt_m1;
  W(PM_Vzyt) = Vzyt; W(PM_Vyzt) = Vyzt;
  // STAGE 11
  double V_max, V_min;
  V_max = v[Vypx]; V_min = v[Vynx];
  v[Vynx] = Vxt - Vyxt - V_max;
  v[Vypx] = Vxt + Vyxt - V_min;
  V_max = v[Vzpx]; V_min = v[Vznx];
  v[Vznx] = Vxt + Vzxt - V_max;
  v[Vzpx] = Vxt - Vzxt - V_min;
  V_max = v[Vzpy]; V_min = v[Vzny];
  v[Vzny] = Vyt - Vzyt - V_max;
  v[Vzpy] = Vyt + Vzyt - V_min;
  V_max = v[Vxpy]; V_min = v[Vxny];
  v[Vxny] = Vyt + Vxyt - V_max;
  v[Vxpy] = Vyt - Vxyt - V_min;
  V_max = v[Vxpz]; V_min = v[Vxnz];
  v[Vxnz] = Vzt - Vxzt - V_max;
  v[Vxpz] = Vzt + Vxzt - V_min;
  V_max = v[Vypz]; V_min = v[Vynz];
  v[Vynz] = Vzt + Vyzt - V_max;
  v[Vypz] = Vzt - Vyzt - V_min;
  // capacitive stub
  W(PM_Vosxr) = Vxt - Vosx; W(PM_Vosyr) = Vyt - Vosy; W(PM_Voszr) = Vzt - Vosz;
  W(PM_Vosx_m1) = Vosx_m1; W(PM_Vosx) = Vosx;
  W(PM_Vosy_m1) = Vosy_m1; W(PM_Vosy) = Vosy;
  W(PM_Vosz_m1) = Vosz_m1; W(PM_Vosz) = Vosz;
  // inductive stub
  const double Vsx = (1.0 / Px0) * (Qx0 * Ixt + Qx1 * Ixt_m1 + Qx2 * Ixt_m2 - Px1 * Vsx_m1 - Px2 * Vsx_m2);
  const double Vsy = (1.0 / Py0) * (Qy0 * Iyt + Qy1 * Iyt_m1 + Qy2 * Iyt_m2 - Py1 * Vsy_m1 - Py2 * Vsy_m2);
  const double Vsz = (1.0 / Pz0) * (Qz0 * Izt + Qz1 * Izt_m1 + Qz2 * Izt_m2 - Pz1 * Vsz_m1 - Pz2 * Vsz_m2);
  W(PM_Vssxr) = (Vsx + Vssx); W(PM_Vssyr) = (Vsy + Vssy); W(PM_Vsszr) = (Vsz + Vssz);
  W(PM_Vsx_m1) = Vsx_m1; W(PM_Vsx) = Vsx;
  W(PM_Vsy_m1) = Vsy_m1; W(PM_Vsy) = Vsy;
  W(PM_Vsz_m1) = Vsz_m1; W(PM_Vsz) = Vsz;
  // STAGE 12
  const int orow = d.pm_out[i];
  if (orow >= 0) {
    double* fo = d.ring + ((long long)slot * d.n_out_rows + orow) * 6;
    fo[0] = -Vxt / dl; fo[1] = -Vyt / dl; fo[2] = -Vzt / dl;
    fo[3] = Ixt / dl; fo[4] = Iyt / dl; fo[5] = Izt / dl;
  }
  // STAGE 13
  v[Vynx] = v[Vynx] * exp_y; v[Vypx] = v[Vypx] * exp_y;
  v[Vznx] = v[Vznx] * exp_z; v[Vzpx] = v[Vzpx] * exp_z;
  v[Vxny] = v[Vxny] * exp_x; v[Vxpy] = v[Vxpy] * exp_x;
  v[Vzny] = v[Vzny] * exp_z; v[Vzpy] = v[Vzpy] * exp_z;
  v[Vxnz] = v[Vxnz] * exp_x; v[Vxpz] = v[Vxpz] * exp_x;
  v[Vynz] = v[Vynz] * exp_y; v[Vypz] = v[Vypz] * exp_y;
  store_cell(Lout, d.P, c, v);
}

template <typename T, bool PLAIN, bool PULL>
__global__ __launch_bounds__(256) void k_pml_material(Dev d, const T* __restrict__ Lin, T* __restrict__ Lout,
                                                     PullArgs a, int slot) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < d.n_pm) pml_material_update<T, PLAIN, PULL>(d, Lin, Lout, a, slot, i);
}

// Thevenin node of a special cell (scatter.F90:274-547): materials,
// optional point source and output row, filter update, 12-port scatter.
// IO = false compiles out the excitation and output branches.
template <bool IO>
__device__ inline void node_update(const Dev& d, double v[12], int mnum, long long f, int er, int orow,
                                   int timestep, int slot) {
  const double Z0 = kZ0, dl = d.dl;
  const double Z04 = Z0 * 4.0, Y04 = kY0 * 4.0;
  const tlm_volume_material* vm = mnum ? &d.vmat[mnum - 1] : nullptr;
  const int mtype = vm ? vm->type : 0;
  double VEx = (v[Vx_ymin] + v[Vx_ymax] + v[Vx_zmin] + v[Vx_zmax]) / 2.0;
  double VEy = (v[Vy_xmin] + v[Vy_xmax] + v[Vy_zmin] + v[Vy_zmax]) / 2.0;
  double VEz = (v[Vz_xmin] + v[Vz_xmax] + v[Vz_ymin] + v[Vz_ymax]) / 2.0;
  double Yx = Y04, Yy = Y04, Yz = Y04;
  double VHx = 2.0 * (v[Vz_ymax] - v[Vy_zmax] - v[Vz_ymin] + v[Vy_zmin]);
  double VHy = 2.0 * (v[Vx_zmax] - v[Vz_xmax] - v[Vx_zmin] + v[Vz_xmin]);
  double VHz = 2.0 * (v[Vy_xmax] - v[Vx_ymax] - v[Vy_xmin] + v[Vx_ymin]);
  double Zx = Z04, Zy = Z04, Zz = Z04;
  double Vox = 0.0, Voy = 0.0, Voz = 0.0, Ys = 0.0;
  const long long nf = d.n_filt;
  if (mtype == TLM_VMAT_PEC) {
    VEx = 0.0; VEy = 0.0; VEz = 0.0;
  } else if (mtype == TLM_VMAT_PMC) {
    VHx = 0.0; VHy = 0.0; VHz = 0.0;
  } else if (mtype == TLM_VMAT_DISPERSIVE) {
    if (vm->eps_filter_exists) {
      Ys = vm->Ys_eps_f;
      const int o = vm->eps_order;
      zf_timeshift(d.eps_w, nf, f, o);
      Vox = zf_evaluate(vm->eps_a, vm->eps_b, d.eps_w, nf, f, o, 0.0);
      zf_timeshift(d.eps_w, nf, f + 1, o);
      Voy = zf_evaluate(vm->eps_a, vm->eps_b, d.eps_w, nf, f + 1, o, 0.0);
      zf_timeshift(d.eps_w, nf, f + 2, o);
      Voz = zf_evaluate(vm->eps_a, vm->eps_b, d.eps_w, nf, f + 2, o, 0.0);
      // add_shunt_admittance (scatter.F90:606-626)
      VEx = (VEx * Yx + Vox * Ys) / (Yx + Ys); Yx = Yx + Ys;
      VEy = (VEy * Yy + Voy * Ys) / (Yy + Ys); Yy = Yy + Ys;
      VEz = (VEz * Yz + Voz * Ys) / (Yz + Ys); Yz = Yz + Ys;
    }
    const double Ge = vm->Ge;
    VEx = (VEx * Yx + 0.0 * Ge) / (Yx + Ge); Yx = Yx + Ge;
    VEy = (VEy * Yy + 0.0 * Ge) / (Yy + Ge); Yy = Yy + Ge;
    VEz = (VEz * Yz + 0.0 * Ge) / (Yz + Ge); Yz = Yz + Ge;
    if (vm->mu_filter_exists) {
      const double Zs = vm->Zs_mu_f;
      const int o = vm->mu_order;
      zf_timeshift(d.mu_w, nf, f, o);
      const double Vsx = zf_evaluate(vm->mu_a, vm->mu_b, d.mu_w, nf, f, o, 0.0);
      zf_timeshift(d.mu_w, nf, f + 1, o);
      const double Vsy = zf_evaluate(vm->mu_a, vm->mu_b, d.mu_w, nf, f + 1, o, 0.0);
      zf_timeshift(d.mu_w, nf, f + 2, o);
      const double Vsz = zf_evaluate(vm->mu_a, vm->mu_b, d.mu_w, nf, f + 2, o, 0.0);
      // add_series_impedance (scatter.F90:641-661)
      VHx = VHx + Vsx; Zx = Zx + Zs;
      VHy = VHy + Vsy; Zy = Zy + Zs;
      VHz = VHz + Vsz; Zz = Zz + Zs;
    }
    const double Rm = vm->Rm;
    VHx = VHx + 0.0; Zx = Zx + Rm;
    VHy = VHy + 0.0; Zy = Zy + Rm;
    VHz = VHz + 0.0; Zz = Zz + Rm;
  }
  double Vx = VEx, Vy = VEy, Vz = VEz;
  double Ix = VHx / Zx, Iy = VHy / Zy, Iz = VHz / Zz;
  if (IO && er >= 0) {
    // cell_excitation_field(row,1:9) at this timestep (excitation_points.F90:262-275)
    double fl[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int s = d.cex_start[er]; s < d.cex_start[er + 1]; s++) {
      const tlm_point_source& p = d.ps[d.cex_src[s]];
      double val = d.value[(long long)(p.function - 1) * d.n_timesteps + (timestep - 1)];
      fl[p.component - 1] = fl[p.component - 1] + val;
    }
    const int32_t* h = d.cex_type + 6 * (long long)er;
    const double Isx = fl[6] * dl * dl, Isy = fl[7] * dl * dl, Isz = fl[8] * dl * dl;
    Vx = (double)h[0] * Vx - fl[0] * dl / 2.0 + Isx / Yx;
    Vy = (double)h[1] * Vy - fl[1] * dl / 2.0 + Isy / Yy;
    Vz = (double)h[2] * Vz - fl[2] * dl / 2.0 + Isz / Yz;
    Ix = (double)h[3] * Ix + fl[3] * dl / 2.0;
    Iy = (double)h[4] * Iy + fl[4] * dl / 2.0;
    Iz = (double)h[5] * Iz + fl[5] * dl / 2.0;
  }
  if (IO && orow >= 0) {
    double* fo = d.ring + ((long long)slot * d.n_out_rows + orow) * 6;
    fo[0] = -Vx / dl; fo[1] = -Vy / dl; fo[2] = -Vz / dl;
    fo[3] = Ix / dl; fo[4] = Iy / dl; fo[5] = Iz / dl;
  }
  if (mtype == TLM_VMAT_DISPERSIVE) {
    if (vm->eps_filter_exists) {
      const double Ysx = vm->Ys_eps_f;
      const int o = vm->eps_order;
      zf_evaluate(vm->eps_a, vm->eps_b, d.eps_w, nf, f, o, (Vx - Vox) * Ysx);
      zf_evaluate(vm->eps_a, vm->eps_b, d.eps_w, nf, f + 1, o, (Vy - Voy) * Ysx);
      zf_evaluate(vm->eps_a, vm->eps_b, d.eps_w, nf, f + 2, o, (Vz - Voz) * Ysx);
    }
    if (vm->mu_filter_exists) {
      const int o = vm->mu_order;
      zf_evaluate(vm->mu_a, vm->mu_b, d.mu_w, nf, f, o, -Ix);
      zf_evaluate(vm->mu_a, vm->mu_b, d.mu_w, nf, f + 1, o, -Iy);
      zf_evaluate(vm->mu_a, vm->mu_b, d.mu_w, nf, f + 2, o, -Iz);
    }
  }
  scatter_ports(v, Vx, Vy, Vz, Ix, Iy, Iz);
}

// node_update (IO false) of a dispersive cell whose filters have order <= 1,
// with the filter words in registers: e0/m0 hold each component's word 0
// (loaded by the caller ahead of the link pull), and the new words 0 and 1
// are stored at the end.  The operations are node_update's, zf_timeshift
// and zf_evaluate's in the same order (the shift moves word 0 to word 1;
// word 1 is never read before it is overwritten), so the results are the
// same bits.
__device__ inline void zf1_words(const double* a, const double* b, int order, double w1, double G, double& W0,
                                 double& F0) {
  W0 = G;
  if (order >= 1) W0 = W0 - b[1] * w1;
  W0 = W0 / b[0];
  F0 = 0.0;
  F0 = F0 + a[0] * W0;
  if (order >= 1) F0 = F0 + a[1] * w1;
}

__device__ inline void mat_node_low(const Dev& d, double v[12], const tlm_volume_material& vm, long long f,
                                    const double e0[3], const double m0[3]) {
  const double Z04 = kZ0 * 4.0, Y04 = kY0 * 4.0;
  double VE[3] = {(v[Vx_ymin] + v[Vx_ymax] + v[Vx_zmin] + v[Vx_zmax]) / 2.0,
                  (v[Vy_xmin] + v[Vy_xmax] + v[Vy_zmin] + v[Vy_zmax]) / 2.0,
                  (v[Vz_xmin] + v[Vz_xmax] + v[Vz_ymin] + v[Vz_ymax]) / 2.0};
  double VH[3] = {2.0 * (v[Vz_ymax] - v[Vy_zmax] - v[Vz_ymin] + v[Vy_zmin]),
                  2.0 * (v[Vx_zmax] - v[Vz_xmax] - v[Vx_zmin] + v[Vz_xmin]),
                  2.0 * (v[Vy_xmax] - v[Vx_ymax] - v[Vy_xmin] + v[Vx_ymin])};
  double Y[3] = {Y04, Y04, Y04}, Z[3] = {Z04, Z04, Z04}, Vo[3] = {0.0, 0.0, 0.0};
  const long long nf = d.n_filt;
  const bool fe = vm.eps_filter_exists, fm = vm.mu_filter_exists;
  const int oe = vm.eps_order, om = vm.mu_order;
  if (fe) {
    const double Ys = vm.Ys_eps_f;
#pragma unroll
    for (int c = 0; c < 3; c++) {
      double W0;
      zf1_words(vm.eps_a, vm.eps_b, oe, e0[c], 0.0, W0, Vo[c]);
      VE[c] = (VE[c] * Y[c] + Vo[c] * Ys) / (Y[c] + Ys);
      Y[c] = Y[c] + Ys;
    }
  }
  const double Ge = vm.Ge;
#pragma unroll
  for (int c = 0; c < 3; c++) {
    VE[c] = (VE[c] * Y[c] + 0.0 * Ge) / (Y[c] + Ge);
    Y[c] = Y[c] + Ge;
  }
  if (fm) {
    const double Zs = vm.Zs_mu_f;
#pragma unroll
    for (int c = 0; c < 3; c++) {
      double W0, Vs;
      zf1_words(vm.mu_a, vm.mu_b, om, m0[c], 0.0, W0, Vs);
      VH[c] = VH[c] + Vs;
      Z[c] = Z[c] + Zs;
    }
  }
  const double Rm = vm.Rm;
#pragma unroll
  for (int c = 0; c < 3; c++) {
    VH[c] = VH[c] + 0.0;
    Z[c] = Z[c] + Rm;
  }
  const double Vx = VE[0], Vy = VE[1], Vz = VE[2];
  const double Ix = VH[0] / Z[0], Iy = VH[1] / Z[1], Iz = VH[2] / Z[2];
  if (fe) {
    const double Ysx = vm.Ys_eps_f;
    const double G[3] = {(Vx - Vo[0]) * Ysx, (Vy - Vo[1]) * Ysx, (Vz - Vo[2]) * Ysx};
#pragma unroll
    for (int c = 0; c < 3; c++) {
      double W0, F0;
      zf1_words(vm.eps_a, vm.eps_b, oe, e0[c], G[c], W0, F0);
      d.eps_w[f + c] = W0;
      if (oe >= 1) d.eps_w[nf + f + c] = e0[c];
    }
  }
  if (fm) {
    const double G[3] = {-Ix, -Iy, -Iz};
#pragma unroll
    for (int c = 0; c < 3; c++) {
      double W0, F0;
      zf1_words(vm.mu_a, vm.mu_b, om, m0[c], G[c], W0, F0);
      d.mu_w[f + c] = W0;
      if (om >= 1) d.mu_w[nf + f + c] = m0[c];
    }
  }
  scatter_ports(v, Vx, Vy, Vz, Ix, Iy, Iz);
}

// gen_code: the reference's update code, 0 for a free-space cell beside a
// special face, negative for a host-path cell (left incident: the host
// scatters it between tlm_scatter and tlm_connect).
template <bool PLAIN>
__device__ inline void generic_node(const Dev& d, double v[12], int timestep, int slot, int i, int code) {
  if (code == 0) {
    if (PLAIN) scatter_free_plain(v); else scatter_free_fast(v);
  } else if (code > 0) {
    node_update<true>(d, v, d.gen_mat[i], d.gen_filt[i], d.gen_exc[i], d.gen_out[i], timestep, slot);
  }
}
template <bool PLAIN>
__device__ inline void generic_node(const Dev& d, double v[12], int timestep, int slot, int i) {
  generic_node<PLAIN>(d, v, timestep, slot, i, d.gen_code[i]);
}

// Generic special cells: point sources, probes, cells touching a special
// face; the Thevenin node, or the free-space branch for code 0 cells that
// only touch a special face.
// FREE = true: a cell of code 0 (free space beside a special face), whose
// node is the free-space scatter; the Thevenin node is not compiled in.
template <typename T, bool PLAIN, bool PULL, bool FREE = false>
__device__ inline void generic_update(const Dev& d, const T* __restrict__ Lin, T* __restrict__ Lout,
                                      const PullArgs& a, int timestep, int slot, int i) {
  const long long c = d.gen_cell[i];
  const int code = FREE ? 0 : d.gen_code[i];  // read with the list entry, not after the pull
  double v[12];
  if (PULL) {
    int x, y, z, sg;
    cell_coords(d, c, x, y, z, sg);
    pull_cell<T, PLAIN>(d, Lin, c, x, y, z, sg, d.gen_fd[i], a, v);
  } else {
#pragma unroll
    for (int p = 0; p < 12; p++) v[p] = ld(Lin, d.P, p, c);
  }
  flush_small(v, a.small);
  if (FREE) {
    if (PLAIN) scatter_free_plain(v); else scatter_free_fast(v);
  } else {
    generic_node<PLAIN>(d, v, timestep, slot, i, code);
  }
  store_cell(Lout, d.P, c, v);
}

// Generic list entries [i0, i1).
template <typename T, bool PLAIN, bool PULL, bool FREE>
__global__ __launch_bounds__(256) void k_generic(Dev d, const T* __restrict__ Lin, T* __restrict__ Lout,
                                                PullArgs a, int timestep, int slot, int i0, int i1) {
  const int i = i0 + blockIdx.x * blockDim.x + threadIdx.x;
  if (i < i1) generic_update<T, PLAIN, PULL, FREE>(d, Lin, Lout, a, timestep, slot, i);
}

// Material cells (dispersive, PEC, PMC) without sources, probes or special
// faces: the bulk of a dielectric body, with compact per-cell metadata.
template <typename T, bool PLAIN, bool PULL>
__device__ inline void material_update(const Dev& d, const T* __restrict__ Lin, T* __restrict__ Lout,
                                       const PullArgs& a, int i) {
  const long long c = d.mat_cell[i];
  double v[12];
  if (PULL) {
    int x, y, z, sg;
    cell_coords(d, c, x, y, z, sg);
    pull_cell<T, PLAIN>(d, Lin, c, x, y, z, sg, -1, a, v);
  } else {
#pragma unroll
    for (int p = 0; p < 12; p++) v[p] = ld(Lin, d.P, p, c);
  }
  flush_small(v, a.small);
  node_update<false>(d, v, d.mat_num[i], d.mat_filt[i], -1, -1, 0, 0);
  store_cell(Lout, d.P, c, v);
}

template <typename T, bool PLAIN, bool PULL>
__global__ __launch_bounds__(256) void k_material(Dev d, const T* __restrict__ Lin, T* __restrict__ Lout,
                                                 PullArgs a, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) material_update<T, PLAIN, PULL>(d, Lin, Lout, a, i);
}

// Pure bulk lines: workgroup b = storage cells [n b, n b + n) (n = 64 for
// the direct line kernel, 256 for the block tables); cells
// of lines with any special cell are left to the mixed-line and class
// kernels, padding cells to no one.
template <typename T, bool PLAIN, bool PULL>
__device__ inline void bulk_line_cell(const Dev& d, const T* __restrict__ Lin, T* __restrict__ Lout,
                                      const PullArgs& a, long long c) {
  if (c >= d.seg_base[d.nseg]) return;
  if (d.line_info[c / d.line_len] != 0u) return;
  int x, y, z, sg;
  if (!cell_coords(d, c, x, y, z, sg)) return;
  bulk_update<T, PLAIN, PULL>(d, Lin, Lout, a, c, x, y, z, sg);
}

// Pure bulk 64-cell storage chunks, one wavefront each.  The chunk table
// holds, per chunk with at least one pure bulk line, a record that every
// lane of the wavefront reads at the same address (a scalar load):
//   x = chunk << 4 | mask of its pure bulk lines (bit q: lanes q*L..q*L+L-1)
//   y = x of its first cell | segment << 16,  z = y,  w = z of that cell.
// A lane's coordinates follow from the first cell's by at most one row wrap
// in a segment at least 64 cells wide; no per-cell table word or division
// precedes the link loads.  (Reading each line's class from the line table
// before the loads measured 7% slower: 0.240 against 0.224 ms on the 256^3
// workload.)  mode 1: each XCD takes a contiguous range of table records.
// bulk_update of cell c0 + k of a chunk (c0 the same in every lane) with
// the chunk's addressing.
template <typename T, bool PLAIN, bool PULL>
__device__ inline void bulk_fast_cell(const Dev& d, const T* __restrict__ Lin, T* __restrict__ Lout,
                                      const PullArgs& a, long long c0, uint32_t k, int x, int y, int z, int sg) {
  double v[12];
  if (PULL) {
    chunk_pull<T, PLAIN>(d, Lin, a, c0, k, x, y, z, sg, v);
  } else {
#pragma unroll
    for (int p = 0; p < 12; p++) v[p] = (double)lk_get_u(Lin, d.P, p, c0, k);
  }
  flush_small(v, a.small);
  if (PLAIN) scatter_free_plain(v); else scatter_free_fast(v);
  store_links_u(Lout, d.P, c0, k, v);
}

template <typename T, bool PLAIN, bool PULL>
__device__ inline void bulk_chunk_cell(const Dev& d, const T* __restrict__ Lin, T* __restrict__ Lout,
                                       const PullArgs& a, int4 e, int lane) {
  if (!((e.x >> (lane / d.line_len)) & 1)) return;
  const long long c = (long long)(e.x >> 4) * kChunk + lane;
  int sg = e.y >> 16, x = (e.y & 0xffff) + lane, y = e.z, z = e.w;
  const int w = seg_width(d, sg);
  if (w >= kChunk) {
    if (x >= d.seg_x0[sg + 1]) {
      x -= w;
      if (++y == d.ny) { y = 0; z++; }
    }
    if (z >= d.nz) return;  // padding after the segment's last row
  } else if (!cell_coords(d, c, x, y, z, sg)) {
    return;
  }
  bulk_fast_cell<T, PLAIN, PULL>(d, Lin, Lout, a, (long long)(e.x >> 4) * kChunk, (uint32_t)lane, x, y, z, sg);
}

template <typename T, bool PLAIN, bool PULL>
__global__ __launch_bounds__(kChunk) void k_bulk_chunks(Dev d, const T* __restrict__ Lin, T* __restrict__ Lout,
                                                        PullArgs a, const int4* __restrict__ tab, int mode) {
  bulk_chunk_cell<T, PLAIN, PULL>(d, Lin, Lout, a, tab[xcd_block(blockIdx.x, gridDim.x, mode)], threadIdx.x);
}

// Material lines: lines holding free-space cells and material cells of one
// material number (a dielectric body's interior and its surface rows).
// One wavefront per 64-cell chunk and material number, from two records
// read by all lanes at one address (scalar loads):
//   e = (chunk << 4 | mask of the chunk's lines it writes,
//        x of the chunk's first cell | segment << 16, y, z)
//   f = (material number, lanes that are material cells (64 bits),
//        filter slot of the first of them)
// The material cells of a chunk hold consecutive filter slots (3 each), so
// a lane's slot follows from the count of material lanes below it.  Every
// cell has six ordinary faces: the bulk pull serves both kinds.  The node
// updates are the reference's: scatter.F90:222-272 for free space, the
// Thevenin node (:274-548) for the material cells.
// LOW = true: records of materials whose filters have order <= 1 only
// (mat_node_low; the general node update is not compiled in: 78 against
// 100 registers).
template <typename T, bool PLAIN, bool PULL, bool LOW = false>
__device__ inline void mat_line_cell(const Dev& d, const T* __restrict__ Lin, T* __restrict__ Lout,
                                     const PullArgs& a, int4 e, int4 f, int lane) {
  if (!((e.x >> (lane / d.line_len)) & 1)) return;
  const long long c = (long long)(e.x >> 4) * kChunk + lane;
  int sg = e.y >> 16, x = (e.y & 0xffff) + lane, y = e.z, z = e.w;
  const int w = seg_width(d, sg);
  if (w >= kChunk) {
    if (x >= d.seg_x0[sg + 1]) {
      x -= w;
      if (++y == d.ny) { y = 0; z++; }
    }
    if (z >= d.nz) return;  // padding after the segment's last row
  } else if (!cell_coords(d, c, x, y, z, sg)) {
    return;
  }
  const uint64_t lanes = (uint64_t)(uint32_t)f.y | ((uint64_t)(uint32_t)f.z << 32);
  // a material lane's filter words (one per component and filter) are
  // loaded here, so that they are in flight with the link pull
  const bool matl = (lanes >> lane) & 1u;
  const tlm_volume_material& vm = d.vmat[f.x > 0 ? f.x - 1 : 0];
  const bool low = LOW || (vm.type == TLM_VMAT_DISPERSIVE && (!vm.eps_filter_exists || vm.eps_order <= 1) &&
                           (!vm.mu_filter_exists || vm.mu_order <= 1));
  const long long slot = matl ? f.w + 3 * (long long)__popcll(lanes & ((1ull << lane) - 1ull)) : 0;
  double e0[3] = {0.0, 0.0, 0.0}, m0[3] = {0.0, 0.0, 0.0};
  if (matl && low) {
    if (vm.eps_filter_exists) {
#pragma unroll
      for (int k = 0; k < 3; k++) e0[k] = d.eps_w[slot + k];
    }
    if (vm.mu_filter_exists) {
#pragma unroll
      for (int k = 0; k < 3; k++) m0[k] = d.mu_w[slot + k];
    }
  }
  double v[12];
  const long long c0 = (long long)(e.x >> 4) * kChunk;
  if (PULL) {
    chunk_pull<T, PLAIN>(d, Lin, a, c0, (uint32_t)lane, x, y, z, sg, v);
  } else {
#pragma unroll
    for (int p = 0; p < 12; p++) v[p] = (double)lk_get_u(Lin, d.P, p, c0, (uint32_t)lane);
  }
  flush_small(v, a.small);
  if (matl) {
    if (LOW || low) mat_node_low(d, v, vm, slot, e0, m0);
    else node_update<false>(d, v, f.x, slot, -1, -1, 0, 0);
  } else if (PLAIN) {
    scatter_free_plain(v);
  } else {
    scatter_free_fast(v);
  }
  store_links_u(Lout, d.P, c0, (uint32_t)lane, v);
}

template <typename T, bool PLAIN, bool PULL, bool LOW>
__global__ __launch_bounds__(kChunk) void k_mat_lines(Dev d, const T* __restrict__ Lin, T* __restrict__ Lout,
                                                      PullArgs a, const int4* __restrict__ rec) {
  mat_line_cell<T, PLAIN, PULL, LOW>(d, Lin, Lout, a, rec[2 * blockIdx.x], rec[2 * blockIdx.x + 1], threadIdx.x);
}

// blocks[0..n_blocks) lists the storage blocks to update (-1: none); a
// workgroup takes per_wg of them in turn: consecutive entries (mode 0, 1),
// or entries 8 apart from its XCD's column (mode 2, the y-band table).
template <typename T, bool PLAIN, bool PULL>
__global__ __launch_bounds__(256) void k_bulk_blocks(Dev d, const T* __restrict__ Lin, T* __restrict__ Lout,
                                                    PullArgs a, const int32_t* __restrict__ blocks, int n_blocks,
                                                    int per_wg, int mode) {
  const int g = blockIdx.x;
  int base, stride;
  if (mode == 2) {
    base = 8 * (g >> 3) * per_wg + (g & 7);
    stride = 8;
  } else {
    base = xcd_block(g, gridDim.x, mode) * per_wg;
    stride = 1;
  }
  for (int k = 0; k < per_wg; k++) {
    const int idx = base + k * stride;
    if (idx >= n_blocks) break;
    const int b = blocks[idx];
    if (b >= 0) bulk_line_cell<T, PLAIN, PULL>(d, Lin, Lout, a, (long long)b * 256 + threadIdx.x);
  }
}

// kMixedBulkPerLane bulk cells per lane of the mixed-line kernel: their
// loads are issued together, so a wavefront of bulk lanes keeps as many
// bytes in flight as the heavier PML lanes at the same occupancy.
#ifndef TLM_MIXED_BULK_PER_LANE
#define TLM_MIXED_BULK_PER_LANE 4
#endif
constexpr int kMixedBulkPerLane = TLM_MIXED_BULK_PER_LANE;

template <typename T, bool PLAIN, bool PULL, int B>
__device__ inline void bulk_update_n(const Dev& d, const T* __restrict__ Lin, T* __restrict__ Lout,
                                     const PullArgs& a, const long long (&cs)[B]) {
  const long long P = d.P;
  T o[B][12], nb[B][12];
  int x[B], y[B], z[B];
#pragma unroll
  for (int b = 0; b < B; b++) {
    x[b] = y[b] = z[b] = 0;
    const long long c = cs[b];
    if (c < 0) continue;
    int sg;
    cell_coords(d, c, x[b], y[b], z[b], sg);
    if (PULL) bulk_own<PLAIN>(d, Lin, P, c, y[b], z[b], o[b]);
    else {
#pragma unroll
      for (int p = 0; p < 12; p++) o[b][p] = Lin[p * P + c];
    }
    if (PULL) {
      const long long sy = seg_width(d, sg), sz = sy * d.ny;
#pragma unroll
      for (int k = 0; k < 12; k++) nb[b][k] = (T)0;
      if (x[b] > 0) {
        const long long n = nb_xm(d, c, x[b], y[b], z[b], sg);
        nb[b][0] = Lin[Vy_xmax * P + n]; nb[b][1] = Lin[Vz_xmax * P + n];
      }
      if (x[b] + 1 < d.nx) {
        const long long n = nb_xp(d, c, x[b], y[b], z[b], sg);
        nb[b][2] = Lin[Vy_xmin * P + n]; nb[b][3] = Lin[Vz_xmin * P + n];
      }
      if (y[b] > 0) { nb[b][4] = Lin[Vx_ymax * P + c - sy]; nb[b][5] = Lin[Vz_ymax * P + c - sy]; }
      if (y[b] + 1 < d.ny) { nb[b][6] = Lin[Vx_ymin * P + c + sy]; nb[b][7] = Lin[Vz_ymin * P + c + sy]; }
      if (z[b] > 0) { nb[b][8] = Lin[Vx_zmax * P + c - sz]; nb[b][9] = Lin[Vy_zmax * P + c - sz]; }
      if (z[b] + 1 < d.nz) { nb[b][10] = Lin[Vx_zmin * P + c + sz]; nb[b][11] = Lin[Vy_zmin * P + c + sz]; }
    }
  }
#pragma unroll
  for (int b = 0; b < B; b++) {
    const long long c = cs[b];
    if (c < 0) continue;
    double v[12];
    if (PULL) {
      double od[12];
#pragma unroll
      for (int p = 0; p < 12; p++) od[p] = (double)o[b][p];
      bulk_incident<PLAIN>(d, a, x[b], y[b], z[b], od, nb[b][0], nb[b][1], nb[b][2], nb[b][3], nb[b][4], nb[b][5],
                           nb[b][6], nb[b][7], nb[b][8], nb[b][9], nb[b][10], nb[b][11], v);
    } else {
#pragma unroll
      for (int p = 0; p < 12; p++) v[p] = (double)o[b][p];
    }
    flush_small(v, a.small);
    if (PLAIN) scatter_free_plain(v); else scatter_free_fast(v);
    store_cell(Lout, P, c, v);
  }
}

// Mixed lines: a workgroup owns whole 128-byte lines (as many as fit in
// 256 lanes with kMixedBulkPerLane bulk cells per lane); its lanes hold the
// special cells sorted by class (PML, generic, material) and then the bulk
// cells, so each wavefront mostly runs one class and every line is
// completed by one workgroup (one L2) within the launch.
//   slot_a[lane] = cell | class << 28 (-1: idle lane)
//   slot_b[k][lane] = class-list position (special), or the k-th further
//                     bulk cell (-1: none)
template <typename T, bool PLAIN, bool PULL>
__global__ __launch_bounds__(256) void k_mixed_lines(Dev d, const T* __restrict__ Lin, T* __restrict__ Lout,
                                                    PullArgs a, const int32_t* __restrict__ slot_a,
                                                    const int32_t* __restrict__ slot_b, long long n_slots,
                                                    int timestep, int slot) {
  const long long s = (long long)blockIdx.x * 256 + threadIdx.x;
  const int32_t w = slot_a[s];
  if (w < 0) return;
  const long long c = w & kCellMask;
  const int cls = (w >> 28) & 3;
  if (cls == 0) {
    long long cs[kMixedBulkPerLane];
    cs[0] = c;
#pragma unroll
    for (int k = 1; k < kMixedBulkPerLane; k++) cs[k] = slot_b[(k - 1) * n_slots + s];
    bulk_update_n<T, PLAIN, PULL, kMixedBulkPerLane>(d, Lin, Lout, a, cs);
    return;
  }
  const int i = slot_b[s];
  if (cls == 1) pml_update<T, PLAIN, PULL>(d, Lin, Lout, a, slot, i);
  else if (cls == 2) material_update<T, PLAIN, PULL>(d, Lin, Lout, a, i);
  else generic_update<T, PLAIN, PULL>(d, Lin, Lout, a, timestep, slot, i);
}

// Node updates on incident voltages already in v (no loads or stores of
// links): the material and generic cells' part of material_update and
// generic_update.
__device__ inline void material_node(const Dev& d, double v[12], int i) {
  node_update<false>(d, v, d.mat_num[i], d.mat_filt[i], -1, -1, 0, 0);
}

// Cells (lanes) per workgroup of the LDS-staged mixed-line kernel.
#ifndef TLM_LDS_CELLS
#define TLM_LDS_CELLS 256
#endif
constexpr int kLdsCells = TLM_LDS_CELLS;

// Mixed lines, LDS-staged.  A workgroup owns kLdsCells/L whole lines.
//   phase 1, cell order: lane t pulls cell t's incident voltages (plane
//            loads coalesced along the lines) into LDS;
//   phase 2, class order: lane t runs the node update of the t-th entry of
//            the workgroup's class-sorted slot list (PML, generic, material,
//            bulk) on its LDS values;
//   phase 3, cell order: lane t stores cell t's scattered voltages, so every
//            line is written whole and coalesced.
// cell_fd: face descriptor of each cell position (-1 none; -2 no cell here:
// padding, or a cell of another kernel).  slot_p: class << 16 | position,
// or -1; slot_i: the class-list position of special cells.  CM: the special
// classes the workgroups of this launch hold (bit 0 PML, 1 material, 2
// generic); the paths of absent classes are compiled out, which lowers the
// register count (the PML path sets it).
template <typename T, bool PLAIN, bool PULL, int CM>
__device__ inline void mixed_lds_body(const Dev& d, const T* __restrict__ Lin, T* __restrict__ Lout,
                                      const PullArgs& a, long long wg, const int32_t* __restrict__ wg_lines,
                                      const int32_t* __restrict__ cell_fd, const int32_t* __restrict__ slot_p,
                                      const int32_t* __restrict__ slot_i, int timestep, int slot, double* sv) {
  constexpr int C = kLdsCells;
  const int t = threadIdx.x;
  const long long base = wg * C;
  const int L = d.line_len;
  const int line = wg_lines[wg * (C / L) + t / L];
  const int fd = cell_fd[base + t];
  const long long c = (long long)line * L + (t % L);
  const bool mine = fd != -2;
  if (mine) {
    double v[12];
    if (PULL) {
      int x, y, z, sg;
      cell_coords(d, c, x, y, z, sg);
      // only PML and generic cells carry face descriptors
      pull_cell<T, PLAIN, (CM & 5) != 0>(d, Lin, c, x, y, z, sg, fd, a, v);
    } else {
#pragma unroll
      for (int p = 0; p < 12; p++) v[p] = ld(Lin, d.P, p, c);
    }
    flush_small(v, a.small);
#pragma unroll
    for (int k = 0; k < 12; k++) sv[k * C + t] = v[k];
  }
  __syncthreads();
  const int sp = slot_p[base + t];
  if (sp >= 0) {
    const int p = sp & 0xffff;
    const int cls = sp >> 16;
    double v[12];
#pragma unroll
    for (int k = 0; k < 12; k++) v[k] = sv[k * C + p];
    if (cls == 0) {
      if (PLAIN) scatter_free_plain(v); else scatter_free_fast(v);
    } else {
      const int i = slot_i[base + t];
      if ((CM & 1) && cls == 1) pml_node<T, PLAIN>(d, v, slot, i, d.pml_cell[i], a, nullptr);
      else if ((CM & 2) && cls == 2) material_node(d, v, i);
      else if ((CM & 4) && cls == 3) generic_node<PLAIN>(d, v, timestep, slot, i);
    }
#pragma unroll
    for (int k = 0; k < 12; k++) sv[k * C + p] = v[k];
  }
  __syncthreads();
  if (mine) {
    double v[12];
#pragma unroll
    for (int k = 0; k < 12; k++) v[k] = sv[k * C + t];
    store_links(Lout, d.P, c, v);
  }
}

// Workgroup wg0 + blockIdx.x of the LDS-staged tables.
template <typename T, bool PLAIN, bool PULL, int CM>
__global__ __launch_bounds__(kLdsCells) void k_mixed_lds(Dev d, const T* __restrict__ Lin, T* __restrict__ Lout,
                                                  PullArgs a, const int32_t* __restrict__ wg_lines,
                                                  const int32_t* __restrict__ cell_fd,
                                                  const int32_t* __restrict__ slot_p,
                                                  const int32_t* __restrict__ slot_i, int timestep, int slot,
                                                  int wg0) {
  __shared__ double sv[12 * kLdsCells];
  mixed_lds_body<T, PLAIN, PULL, CM>(d, Lin, Lout, a, wg0 + (long long)blockIdx.x, wg_lines, cell_fd, slot_p, slot_i,
                                     timestep, slot, sv);
}

// Class masks of the LDS-staged mixed-line launches: workgroups without PML
// cells (the PML path sets the register count: 106 against 168 VGPRs in the
// precision form), and the rest.
constexpr int kMixedMasks[2] = {6, 7};

// Connect only (no scatter): incident voltages into Lout, used to hand the
// state back after a connect.  Bulk variant over all cells; the class
// lists then rewrite their own cells with special-face pulls.
template <typename T, bool PLAIN>
__global__ __launch_bounds__(256) void k_connect_bulk(Dev d, const T* __restrict__ Lin, T* __restrict__ Lout,
                                                     PullArgs a) {
  int x = blockIdx.x * blockDim.x + threadIdx.x;
  int y = blockIdx.y * blockDim.y + threadIdx.y;
  int z = blockIdx.z;
  if (x >= d.nx || y >= d.ny) return;
  int sg;
  const long long c = storage_index(d, x, y, z, sg);
  double v[12];
  pull_cell<T, PLAIN>(d, Lin, c, x, y, z, sg, -1, a, v);
  store_cell(Lout, d.P, c, v);
}

template <typename T, bool PLAIN>
__global__ __launch_bounds__(256) void k_connect_list(Dev d, const T* __restrict__ Lin, T* __restrict__ Lout,
                                                     PullArgs a, const int32_t* cells, const int32_t* fds, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const long long c = cells[i];
  int x, y, z, sg;
  cell_coords(d, c, x, y, z, sg);
  double v[12];
  pull_cell<T, PLAIN>(d, Lin, c, x, y, z, sg, fds ? fds[i] : -1, a, v);
  store_cell(Lout, d.P, c, v);
}

// Host-path cells: the 12 link values of cell cells[k] <-> buf[12 k + p].
template <typename T>
__global__ void k_gather_host(const T* __restrict__ L, long long P, const int32_t* __restrict__ cells, int n,
                              double* __restrict__ buf) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const long long c = cells[k];
  for (int p = 0; p < 12; p++) buf[12 * (long long)k + p] = (double)lk_get(L, P, p, c);
}

template <typename T>
__global__ void k_scatter_host(T* __restrict__ L, long long P, const int32_t* __restrict__ cells, int n,
                               const double* __restrict__ buf) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const long long c = cells[k];
  for (int p = 0; p < 12; p++) lk_put(L, P, p, c, buf[12 * (long long)k + p]);
}

// The four scattered pulses beside each host-path face (tlm_gather_host_faces):
// (a on the + side cell cp, a on the - side cell cm, b on cp, b on cm).
template <typename T>
__global__ void k_gather_host_faces(const T* __restrict__ L, long long P, const int32_t* __restrict__ cp,
                                    const int32_t* __restrict__ cm, const int32_t* __restrict__ ax, int n,
                                    double* __restrict__ buf) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int a = ax[k];
  const int pa = a == 0 ? Vy_xmin : a == 1 ? Vx_ymin : Vx_zmin;
  const int pb = a == 0 ? Vz_xmin : a == 1 ? Vz_ymin : Vy_zmin;
  const int qa = a == 0 ? Vy_xmax : a == 1 ? Vx_ymax : Vx_zmax;
  const int qb = a == 0 ? Vz_xmax : a == 1 ? Vz_ymax : Vy_zmax;
  buf[4 * (long long)k + 0] = (double)lk_get(L, P, pa, cp[k]);
  buf[4 * (long long)k + 1] = (double)lk_get(L, P, qa, cm[k]);
  buf[4 * (long long)k + 2] = (double)lk_get(L, P, pb, cp[k]);
  buf[4 * (long long)k + 3] = (double)lk_get(L, P, qb, cm[k]);
}

// Probe DFT: acc(p,k) += v_p(n) exp(-j 2 pi f_k t_n) dt, t_n = (n-1) dt.
static __global__ void k_dft(const double* __restrict__ ring, int n_out_rows, int slot, int timestep,
                      const int32_t* __restrict__ prow, const int32_t* __restrict__ pcomp, int n_points,
                      const double* __restrict__ freq, int n_freq, double dt, double* __restrict__ acc) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_points * n_freq) return;
  int p = i / n_freq, k = i - p * n_freq;
  double v = ring[((long long)slot * n_out_rows + prow[p]) * 6 + pcomp[p]];
  double time = (double)(timestep - 1) * dt;
  double ph = 2.0 * kPI * freq[k] * time;
  double s, co;
  sincos(ph, &s, &co);
  acc[2 * i] += v * co * dt;
  acc[2 * i + 1] -= v * s * dt;
}

// Frequency-domain outputs (frequency_output_volumes.F90:309-386,
// frequency_output_surfaces.F90:297-421): one lane per item of every output
// of the launch's kind (1 volume, 2 surface) due at this timestep.
// src: the step's cell-output rows (volumes) or the connect's face-output
// block (surfaces), rows wide; acc gains value * exp(j y), y = -2 pi f time.
// One accumulated frequency output (kind 1: cell rows, 2: face rows).
// comp 1..6: a field; 7..9: J = n x H; 10..12: M = -n x E.  Each due step
// adds value * exp(-j 2 pi f time) * scale (scale: 1, or dt for far-field
// surfaces).
struct FreqOut {
  double frequency, scale;
  int kind, comp, first, last, interval;
};
struct FreqItem {
  int out, row, side, normal;
};

static __global__ void k_freq_out(const FreqOut* __restrict__ outs, const FreqItem* __restrict__ items, int n,
                                  int kind, const double* __restrict__ src, int rows, int timestep, double time,
                                  double* __restrict__ acc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const FreqItem it = items[i];
  const FreqOut o = outs[it.out];
  if (o.kind != kind || timestep < o.first || timestep > o.last || (timestep - o.first) % o.interval != 0) return;
  double value;
  if (kind == 1) {
    value = src[(long long)(it.row - 1) * 6 + (o.comp - 1)];
  } else {
    const double* f = src + ((long long)(it.row - 1) * 2 + (it.side - 1)) * 6;
    if (o.comp <= 6) {
      value = f[o.comp - 1];
    } else {
      // J = n x H with the face's outward normal (normx, normy, normz)
      const double nx = it.normal == 1 ? 1.0 : it.normal == -1 ? -1.0 : 0.0;
      const double ny = it.normal == 2 ? 1.0 : it.normal == -2 ? -1.0 : 0.0;
      const double nz = it.normal == 3 ? 1.0 : it.normal == -3 ? -1.0 : 0.0;
      if (o.comp == 7) value = ny * f[5] - nz * f[4];
      else if (o.comp == 8) value = nz * f[3] - nx * f[5];
      else if (o.comp == 9) value = nx * f[4] - ny * f[3];
      // M = -n x E
      else if (o.comp == 10) value = -(ny * f[2] - nz * f[1]);
      else if (o.comp == 11) value = -(nz * f[0] - nx * f[2]);
      else value = -(nx * f[1] - ny * f[0]);
    }
  }
  const double y = -(2.0 * kPI * o.frequency * time);
  double sy, cy;
  sincos(y, &sy, &cy);
  acc[2 * i] += value * (cy * o.scale);
  acc[2 * i + 1] += value * (sy * o.scale);
}

// One frame of a time-domain field output (output_surfaces.F90:660-720,
// output_volumes.F90:535-545): one value per item.  kind 1: the cell row's
// field `comp`; kind 2: the face's field (comp 1..6), J = n x H (7..9),
// |E| (13), |H| (14) or |J| (15), with J taken as the jump between the two
// sides of the face when `both`.
static __global__ void k_field_out(const FreqItem* __restrict__ items, int n, int kind, int comp, int both,
                                   const double* __restrict__ src, double* __restrict__ dst) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const FreqItem it = items[i];
  double value;
  if (kind == 1) {
    value = src[(long long)(it.row - 1) * 6 + (comp - 1)];
  } else {
    const double* f = src + ((long long)(it.row - 1) * 2 + (both ? 0 : it.side - 1)) * 6;
    const double* g = src + ((long long)(it.row - 1) * 2 + 1) * 6;
    const double nx = it.normal == 1 ? 1.0 : it.normal == -1 ? -1.0 : 0.0;
    const double ny = it.normal == 2 ? 1.0 : it.normal == -2 ? -1.0 : 0.0;
    const double nz = it.normal == 3 ? 1.0 : it.normal == -3 ? -1.0 : 0.0;
    // the second side's H: zero unless both
    const double g3 = both ? g[3] : 0.0, g4 = both ? g[4] : 0.0, g5 = both ? g[5] : 0.0;
    const double jx = (ny * f[5] - nz * f[4]) - (ny * g5 - nz * g4);
    const double jy = (nz * f[3] - nx * f[5]) - (nz * g3 - nx * g5);
    const double jz = (nx * f[4] - ny * f[3]) - (nx * g4 - ny * g3);
    if (comp <= 6) value = f[comp - 1];
    else if (comp == 7) value = jx;
    else if (comp == 8) value = jy;
    else if (comp == 9) value = jz;
    else if (comp == 13) value = sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
    else if (comp == 14) value = sqrt(f[3] * f[3] + f[4] * f[4] + f[5] * f[5]);
    else value = sqrt(jx * jx + jy * jy + jz * jz);
  }
  dst[i] = value;
}

// Far-field transform of one surface (far_field_surfaces.F90:370-712):
// one workgroup per angle sums the N and L integrals over the faces and
// forms E_theta, E_phi at r = 1 m.  acc holds the surface's accumulators
// component-major (Jx, Jy, Jz, Mx, My, Mz blocks of nf complex values).
struct FarFieldArgs {
  double beta, dA, theta_min, theta_step, phi_min, phi_step, prop2_re, prop2_im;
  int n_theta, n_phi, dim, nf;
};

__device__ inline void cmul_acc(double& sr, double& si, double ar, double ai, double br, double bi, double s) {
  // s += (a * b) * s_real, complex a, b
  const double pr = ar * br - ai * bi, pi = ar * bi + ai * br;
  sr += pr * s;
  si += pi * s;
}

static __global__ void __launch_bounds__(256) k_far_field(FarFieldArgs s, const double* __restrict__ acc,
                                                          const double* __restrict__ centres,
                                                          double* __restrict__ e) {
  __shared__ double red[8][256];
  const int a = blockIdx.x, tid = threadIdx.x;
  const int it = a / s.n_phi, ip = a - it * s.n_phi;
  const double theta = s.theta_min + (double)it * s.theta_step;
  const double phi = s.phi_min + (double)ip * s.phi_step;
  const double ct = cos(theta), st = sin(theta), cp = cos(phi), sp = sin(phi);
  const double r1 = st * cp, r2 = st * sp, r3 = ct;
  const int nf = s.nf;
  const double* J0 = acc;
  const double* J1 = acc + 2 * (size_t)nf;
  const double* J2 = acc + 4 * (size_t)nf;
  const double* M0 = acc + 6 * (size_t)nf;
  const double* M1 = acc + 8 * (size_t)nf;
  const double* M2 = acc + 10 * (size_t)nf;
  // 3D: N_theta, N_phi, L_theta, L_phi; 2D: N_z, N_phi, L_z, L_phi
  double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int f = tid; f < nf; f += 256) {
    const double* rp = centres + 3 * (size_t)f;
    const double x = s.dim == 2 ? r1 * rp[0] + r2 * rp[1] : r1 * rp[0] + r2 * rp[1] + r3 * rp[2];
    double pim, pre;
    sincos(s.beta * x, &pim, &pre);
    const double jxr = J0[2 * f], jxi = J0[2 * f + 1], jyr = J1[2 * f], jyi = J1[2 * f + 1];
    const double jzr = J2[2 * f], jzi = J2[2 * f + 1];
    const double mxr = M0[2 * f], mxi = M0[2 * f + 1], myr = M1[2 * f], myi = M1[2 * f + 1];
    const double mzr = M2[2 * f], mzi = M2[2 * f + 1];
    if (s.dim == 2) {
      cmul_acc(v[0], v[1], jzr, jzi, pre, pim, s.dA);
      cmul_acc(v[2], v[3], -jxr * sp + jyr * cp, -jxi * sp + jyi * cp, pre, pim, s.dA);
      cmul_acc(v[4], v[5], mzr, mzi, pre, pim, s.dA);
      cmul_acc(v[6], v[7], -mxr * sp + myr * cp, -mxi * sp + myi * cp, pre, pim, s.dA);
    } else {
      cmul_acc(v[0], v[1], jxr * ct * cp + jyr * ct * sp - jzr * st, jxi * ct * cp + jyi * ct * sp - jzi * st, pre,
               pim, s.dA);
      cmul_acc(v[2], v[3], -jxr * sp + jyr * cp, -jxi * sp + jyi * cp, pre, pim, s.dA);
      cmul_acc(v[4], v[5], mxr * ct * cp + myr * ct * sp - mzr * st, mxi * ct * cp + myi * ct * sp - mzi * st, pre,
               pim, s.dA);
      cmul_acc(v[6], v[7], -mxr * sp + myr * cp, -mxi * sp + myi * cp, pre, pim, s.dA);
    }
  }
#pragma unroll
  for (int k = 0; k < 8; k++) red[k][tid] = v[k];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (tid < w) {
#pragma unroll
      for (int k = 0; k < 8; k++) red[k][tid] += red[k][tid + w];
    }
    __syncthreads();
  }
  if (tid == 0) {
    const double nar = red[0][0], nai = red[1][0], npr = red[2][0], npi = red[3][0];
    const double lar = red[4][0], lai = red[5][0], lpr = red[6][0], lpi = red[7][0];
    const double p2r = s.prop2_re, p2i = s.prop2_im;
    double qr, qi, ur, ui;
    if (s.dim == 2) {
      // E_theta = -prop2 (L_phi - Z0 N_z), E_phi = prop2 (L_z + Z0 N_phi)
      qr = lpr - kZ0 * nar; qi = lpi - kZ0 * nai;
      ur = lar + kZ0 * npr; ui = lai + kZ0 * npi;
    } else {
      // E_theta = -prop2 (L_phi + Z0 N_theta), E_phi = prop2 (L_theta - Z0 N_phi)
      qr = lpr + kZ0 * nar; qi = lpi + kZ0 * nai;
      ur = lar - kZ0 * npr; ui = lai - kZ0 * npi;
    }
    e[4 * (size_t)a] = -(p2r * qr - p2i * qi);
    e[4 * (size_t)a + 1] = -(p2r * qi + p2i * qr);
    e[4 * (size_t)a + 2] = p2r * ur - p2i * ui;
    e[4 * (size_t)a + 3] = p2r * ui + p2i * ur;
  }
}

// RCS surface (RCS_surfaces.F90:386-559).  Faces are sorted by d = K.r
// ascending, so the far-field point of a face at one step,
// rcs_point(d), never increases along them and the faces of one point are
// a contiguous run.  One workgroup per point of this step's range sums its
// run (binary search for its ends) and adds the sums to the point.
struct RcsFace {
  int row, side, normal, pad;
  double d;
};
struct RcsArgs {
  double ct, st, cp, sp, dA, time_c0, dl;
  int points_per_cell, offset, nf, p_lo, rows;
};

__host__ __device__ inline int rcs_point(double d, double time_c0, int points_per_cell, double dl, int offset) {
  const double dm = -d + time_c0;
  return (int)((double)points_per_cell * dm / dl) - offset;
}

static __global__ void __launch_bounds__(256) k_rcs(RcsArgs a, const RcsFace* __restrict__ faces,
                                                    const double* __restrict__ src, double* __restrict__ E) {
  __shared__ double red[4][256];
  const int tid = threadIdx.x, p = a.p_lo + blockIdx.x;
  int lo = 0, hi = a.nf;
  while (lo < hi) {  // first face with point <= p
    const int mid = (lo + hi) >> 1;
    if (rcs_point(faces[mid].d, a.time_c0, a.points_per_cell, a.dl, a.offset) <= p) hi = mid;
    else lo = mid + 1;
  }
  const int i0 = lo;
  hi = a.nf;
  while (lo < hi) {  // first face with point < p
    const int mid = (lo + hi) >> 1;
    if (rcs_point(faces[mid].d, a.time_c0, a.points_per_cell, a.dl, a.offset) < p) hi = mid;
    else lo = mid + 1;
  }
  const int i1 = lo;
  const double q = 4.0 * kPI;
  double v[4] = {0, 0, 0, 0};
  for (int i = i0 + tid; i < i1; i += 256) {
    const RcsFace fc = faces[i];
    const double* f = src + ((size_t)(fc.row - 1) * 2 + (fc.side - 1)) * 6;
    const double nx = fc.normal == 1 ? 1.0 : fc.normal == -1 ? -1.0 : 0.0;
    const double ny = fc.normal == 2 ? 1.0 : fc.normal == -2 ? -1.0 : 0.0;
    const double nz = fc.normal == 3 ? 1.0 : fc.normal == -3 ? -1.0 : 0.0;
    const double jx = ny * f[5] - nz * f[4], jy = nz * f[3] - nx * f[5], jz = nx * f[4] - ny * f[3];
    const double mx = -(ny * f[2] - nz * f[1]), my = -(nz * f[0] - nx * f[2]), mz = -(nx * f[1] - ny * f[0]);
    const double nth = (jx * a.ct * a.cp + jy * a.ct * a.sp - jz * a.st) * a.dA / q;
    const double nph = (-jx * a.sp + jy * a.cp) * a.dA / q;
    const double lth = (mx * a.ct * a.cp + my * a.ct * a.sp - mz * a.st) * a.dA / q;
    const double lph = (-mx * a.sp + my * a.cp) * a.dA / q;
    v[0] += -(lph + kZ0 * nth);
    v[1] += lth - kZ0 * nph;
    v[2] += nph - lth / kZ0;
    v[3] += -(nth + lph / kZ0);
  }
#pragma unroll
  for (int k = 0; k < 4; k++) red[k][tid] = v[k];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (tid < w) {
#pragma unroll
      for (int k = 0; k < 4; k++) red[k][tid] += red[k][tid + w];
    }
    __syncthreads();
  }
  if (tid == 0 && i1 > i0) {
#pragma unroll
    for (int k = 0; k < 4; k++) E[4 * (size_t)(p - 1) + k] += red[k][0];
  }
}

// Fortran V(1:12,0:nx+1,0:ny+1,1:nz) <-> planes; thread = storage index.
template <typename T>
__global__ void k_aos_to_planes(Dev d, const double* __restrict__ V, T* __restrict__ L) {
  const long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (c >= d.seg_base[d.nseg]) return;
  int x, y, z, s;
  if (!cell_coords(d, c, x, y, z, s)) return;
  const long long fi =
      12 * ((long long)(x + 1) + (long long)(d.nx + 2) * ((long long)(y + 1) + (long long)(d.ny + 2) * z));
  for (int p = 0; p < 12; p++) lk_put(L, d.P, p, c, V[fi + p]);
}

template <typename T>
__global__ void k_planes_to_aos(Dev d, const T* __restrict__ L, double* __restrict__ V) {
  const long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (c >= d.seg_base[d.nseg]) return;
  int x, y, z, s;
  if (!cell_coords(d, c, x, y, z, s)) return;
  const long long fi =
      12 * ((long long)(x + 1) + (long long)(d.nx + 2) * ((long long)(y + 1) + (long long)(d.ny + 2) * z));
  for (int p = 0; p < 12; p++) V[fi + p] = (double)lk_get(L, d.P, p, c);
}

// k_planes_to_aos for the z-planes [z0, z1) only, into a slab array laid
// out as V(12, 0:nx+1, 0:ny+1, z0:z1-1); one lane per cell of the slab.
template <typename T>
__global__ void k_planes_to_aos_slab(Dev d, const T* __restrict__ L, double* __restrict__ V, int z0, int z1) {
  const long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  const long long plane = (long long)d.nx * d.ny;
  if (k >= plane * (z1 - z0)) return;
  const int x = (int)(k % d.nx);
  const int y = (int)((k / d.nx) % d.ny);
  const int zl = (int)(k / plane);
  int s;
  const long long c = storage_index(d, x, y, z0 + zl, s);
  const long long fi =
      12 * ((long long)(x + 1) + (long long)(d.nx + 2) * ((long long)(y + 1) + (long long)(d.ny + 2) * zl));
  for (int p = 0; p < 12; p++) V[fi + p] = (double)lk_get(L, d.P, p, c);
}

}  // namespace scn
