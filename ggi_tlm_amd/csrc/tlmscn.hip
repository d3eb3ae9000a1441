// tlmscn.hip -- C ABI of the gfx950 SCN scatter/connect engine
// (declarations and reference mapping in include/tlmscn.h).
//
// Host side: validates the reference tables, derives the device layout
// (x-segment storage order, 128-byte line classes, class lists, mixed-line
// workgroup tables, special-face descriptors, excitation CSR tables), owns
// the handle's HIP streams and drives the step kernels of scn_kernels.hpp.
//
// Link-state machine.  The reference keeps one array V that scatter and
// connect update in place.  Here a buffer holds either incident voltages
// (after an upload) or the scattered voltages of the last scatter; connect
// and outer_boundary are recorded as pending and folded into the next
// scatter's pull.  A download materialises the pending connect first, so
// the caller always sees exactly the reference's V.
#include <mutex>

#include "engine.hpp"

using namespace engine;

namespace engine {

// Profiling brackets (no-ops unless enabled).
struct ProfScope {
  tlm_handle* h;
  int cls;
  hipEvent_t a = nullptr, b = nullptr;
  ProfScope(tlm_handle* hh, int c) : h(hh), cls(c) {
    if (!h->profiling) return;
    (void)hipEventCreate(&a);
    (void)hipEventCreate(&b);
    (void)hipEventRecord(a, h->stream);
  }
  ~ProfScope() {
    if (!h->profiling) return;
    (void)hipEventRecord(b, h->stream);
    h->ev.push_back({cls, {a, b}});
  }
};

int resolve_profile(tlm_handle* h) {
  if (h->ev.empty()) return TLM_OK;
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  for (auto& e : h->ev) {
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e.second.first, e.second.second);
    h->prof_ms[e.first] += ms;
    h->prof_launches[e.first] += 1;
    (void)hipEventDestroy(e.second.first);
    (void)hipEventDestroy(e.second.second);
  }
  h->ev.clear();
  return TLM_OK;
}

// Build every device table from the reference tables.
// Filter state bytes a dispersive cell of material m reads and writes per
// step: 3 components x (carried words read and written) for eps and mu, FP64.
static int64_t filter_bytes_of(const tlm_handle* h, int m) {
  if (!m) return 0;
  const tlm_volume_material& vm = h->vmat[m - 1];
  if (vm.type != TLM_VMAT_DISPERSIVE) return 0;
  const int64_t words = (vm.eps_filter_exists ? vm.eps_order + 1 : 0) + (vm.mu_filter_exists ? vm.mu_order + 1 : 0);
  return 3 * 2 * 8 * words;
}

// Surface material types whose face model stays on the host.
static bool host_smat(int type) {
  return type == TLM_SMAT_DISPERSIVE || type == TLM_SMAT_ANISOTROPIC_DISPERSIVE || type == TLM_SMAT_DIODE ||
         type == TLM_SMAT_SPICE || type == TLM_SMAT_SWITCH;
}

// The chunk records keep the chunk number in 27 bits.
static bool k_max_chunks_exceeded(long long storage_cells) {
  return (storage_cells + scn::kChunk - 1) / scn::kChunk >= (1LL << 27);
}

int finalize(tlm_handle* h) {
  if (h->finalized) return TLM_OK;
  if (!h->cells_set) return fail(h, TLM_ERR_STATE, "tlm_set_cell_codes has not been called");
  if (!h->faces_set) return fail(h, TLM_ERR_STATE, "tlm_set_face_codes has not been called");
  const int nx = h->nx, ny = h->ny, nz = h->nz;
  const long long N = h->N;
  int rc;

  // ---- special faces: locate each special face code in scan order -------
  // face_info[code] = (cell index of the face's min-side cell cx, axis)
  std::vector<long long> sf_cell(h->n_special_faces + 1, -1), sf_fnum(h->n_special_faces + 1, -1);
  std::vector<int> sf_axis(h->n_special_faces + 1, -1);
  if (h->n_special_faces > 0) {
    long long fnum = 0;
    int found = 0;
    for (int cz = 0; cz < nz && found < h->n_special_faces; cz++)
      for (int cy = 0; cy < ny && found < h->n_special_faces; cy++)
        for (int cx = 0; cx < nx; cx++) {
          long long c = cx + (long long)nx * (cy + (long long)ny * cz);
          for (int ax = 0; ax < 3; ax++) {
            bool has = (ax == 0 && cx != 0) || (ax == 1 && cy != 0) || (ax == 2 && cz != 0);
            if (!has) continue;
            int code = h->face_code[fnum++];
            if (code != 0) {
              if (code < 1 || code > h->n_special_faces) return fail(h, TLM_ERR_ARG, "face code out of range");
              sf_cell[code] = c;
              sf_axis[code] = ax;
              sf_fnum[code] = fnum - 1;
              found++;
            }
          }
        }
  }
  // surface material type, excitation and output rows per special face
  std::vector<int32_t> sf_mtype(h->n_special_faces + 1, 0), sf_exc(h->n_special_faces + 1, -1);
  std::vector<int32_t> sf_out(h->n_special_faces + 1, -1);
  int n_face_out_rows = 0;
  {
    int er = 0;
    // excitation rows follow the face scan order = code order
    for (int code = 1; code <= h->n_special_faces; code++) {
      int m = h->f_mat[code - 1];
      int mn = m < 0 ? -m : m;
      if (mn) {
        if (mn > (int)h->smat.size()) return fail(h, TLM_ERR_ARG, "surface material number out of range");
        sf_mtype[code] = h->smat[mn - 1];
      }
      if (h->f_exc[code - 1] != 0) sf_exc[code] = er++;
      // output rows follow the face scan order too (output_face_number)
      if (!h->f_out.empty() && h->f_out[code - 1] != 0) sf_out[code] = n_face_out_rows++;
    }
    if (er != h->n_face_rows) return fail(h, TLM_ERR_ARG, "face excitation rows do not match the face codes");
  }
  // host-path faces in the connect's scan order
  std::vector<int32_t> sf_host(h->n_special_faces + 1, -1), hf_codes;
  for (int code = 1; code <= h->n_special_faces; code++)
    if (host_smat(sf_mtype[code])) hf_codes.push_back(code);
  std::sort(hf_codes.begin(), hf_codes.end(), [&](int a, int b) { return sf_fnum[a] < sf_fnum[b]; });
  for (size_t k = 0; k < hf_codes.size(); k++) sf_host[hf_codes[k]] = (int32_t)k;
  // descriptors: per special cell, the special-face codes of its 6 faces
  // (xmin, xmax, ymin, ymax, zmin, zmax).  Key: cell index.
  struct FdEntry { long long cell; int side; int code; };
  std::vector<FdEntry> fe;
  fe.reserve(2 * (size_t)h->n_special_faces);
  for (int code = 1; code <= h->n_special_faces; code++) {
    long long c = sf_cell[code];
    if (c < 0) return fail(h, TLM_ERR_ARG, "special face code not present in face_update_code");
    int ax = sf_axis[code];
    long long off = ax == 0 ? 1 : (ax == 1 ? nx : (long long)nx * ny);
    fe.push_back({c, 2 * ax, code});            // min face of cell c
    fe.push_back({c - off, 2 * ax + 1, code});  // max face of the cell below
  }
  std::sort(fe.begin(), fe.end(), [](const FdEntry& a, const FdEntry& b) { return a.cell < b.cell; });
  std::vector<int32_t> fd;         // 6 per descriptor
  std::vector<long long> fd_cell;  // sorted cells with descriptors
  for (size_t i = 0; i < fe.size();) {
    size_t j = i;
    int32_t codes[6] = {0, 0, 0, 0, 0, 0};
    while (j < fe.size() && fe[j].cell == fe[i].cell) { codes[fe[j].side] = fe[j].code; j++; }
    fd_cell.push_back(fe[i].cell);
    for (int k = 0; k < 6; k++) fd.push_back(codes[k]);
    i = j;
  }

  // ---- cell classes ------------------------------------------------------
  // class per cell: 0 bulk, 1 PML, 2 material, 3 generic, 4 PML holding a material
  struct PmlRec { int32_t c; uint16_t par; int32_t orow, fdc; };
  struct MatRec { int32_t c; uint16_t m; int32_t f; };
  struct GenRec { int32_t c, code, m, f, er, orow, fdc; };
  struct PmRec { int32_t c, par, m, orow, fdc; };  // PML cell holding a material (class 4)
  std::vector<uint8_t> cls(N, 0);  // by reference cell index
  std::vector<PmRec> pmr;
  std::vector<PmlRec> pr;
  std::vector<MatRec> mr;
  std::vector<GenRec> gr;
  long long max_filt = 0;
  int out_rank = 0, exc_rank = 0;
  size_t fdi = 0;  // walks fd_cell in scan order
  int n_pml_state = (int)h->pml_cell_param.size();
  std::vector<int32_t> host_ref;  // host-path cells, reference index, scan order
  for (long long c = 0; c < N; c++) {
    int code = h->cell_code[c];
    bool has_fd = fdi < fd_cell.size() && fd_cell[fdi] == c;
    int32_t fdc = has_fd ? (int32_t)fdi : -1;
    if (has_fd) fdi++;
    if (code == 0 && !has_fd) continue;
    if (code < 0 || code > h->n_special) return fail(h, TLM_ERR_ARG, "cell code out of range");
    int s = code - 1;
    int pmlc = code ? h->c_pml[s] : 0;
    int orow = (code && h->c_out[s]) ? out_rank++ : -1;
    if (code && h->c_cable[s]) {
      // host-path (cable junction) cell: pulled on the device, scattered by
      // the host between tlm_scatter and tlm_connect
      if (pmlc || orow >= 0 || h->c_exc[s])
        return fail(h, TLM_ERR_UNSUPPORTED, "host-path (cable) cells carry no PML, excitation or output row");
      cls[c] = 3;
      gr.push_back({(int32_t)c, -code, 0, 0, -1, -1, fdc});
      host_ref.push_back((int32_t)c);
      continue;
    }
    if (pmlc != 0) {
      if (pmlc < 1 || pmlc > n_pml_state) return fail(h, TLM_ERR_ARG, "PML cell number out of range");
      int m = h->c_mat[s];
      if (m) {
        if (m > (int)h->vmat.size()) return fail(h, TLM_ERR_ARG, "material number out of range");
        int t = h->vmat[m - 1].type;
        if (t == TLM_VMAT_PEC) return fail(h, TLM_ERR_ARG, "ERROR: PEC volume material in PML");
        if (t == TLM_VMAT_PMC) return fail(h, TLM_ERR_ARG, "ERROR: PMC volume material in PML");
      }
      if (h->c_exc[s]) return fail(h, TLM_ERR_ARG, "Excitation intersects PML");
      int par = h->pml_cell_param[pmlc - 1];
      if (par < 1 || par > (int)h->pml_par.size() || par > 65536)
        return fail(h, TLM_ERR_ARG, "PML parameter index out of range");
      if (m && h->vmat[m - 1].type == TLM_VMAT_DISPERSIVE) {
        // set_update_codes.F90:300-333: only order-0 (non-dispersive)
        // S-filters may sit in the PML; their stub Z-filters have order <= 1
        if (h->vmat[m - 1].eps_order > 1 || h->vmat[m - 1].mu_order > 1)
          return fail(h, TLM_ERR_ARG, "PML fatal error: dispersive material found");
        cls[c] = 4;
        pmr.push_back({(int32_t)c, par - 1, m, orow, fdc});
        continue;
      }
      cls[c] = 1;
      pr.push_back({(int32_t)c, (uint16_t)(par - 1), orow, fdc});
    } else {
      int m = code ? h->c_mat[s] : 0;
      int f = code ? h->c_filt[s] : 0;
      if (m) {
        if (m > (int)h->vmat.size()) return fail(h, TLM_ERR_ARG, "material number out of range");
        if (h->vmat[m - 1].type == TLM_VMAT_DISPERSIVE) {
          if (f < 1) return fail(h, TLM_ERR_ARG, "dispersive cell without filter number");
          max_filt = std::max<long long>(max_filt, (long long)f + 3);
        }
      }
      int er = (code && h->c_exc[s]) ? exc_rank++ : -1;
      if (code && m && er < 0 && orow < 0 && fdc < 0 && m <= 65535) {
        cls[c] = 2;
        mr.push_back({(int32_t)c, (uint16_t)m, f});
        continue;
      }
      cls[c] = 3;
      gr.push_back({(int32_t)c, code, m, f, er, orow, fdc});
    }
  }

  // ---- storage layout: x-segments ---------------------------------------------
  // Cut between x and x+1 where at least an eighth of the (y, z) positions
  // change class, keeping the TLM_MAX_SEGMENTS - 1 strongest cuts; segment
  // bases are multiples of 64 cells.
  std::vector<int> seg_x0 = {0};
  if (h->layout == TLM_LAYOUT_X_SEGMENTS && nx > 1) {
    std::vector<long long> T(nx - 1, 0);
    for (long long r = 0; r < (long long)ny * nz; r++) {
      const uint8_t* row = &cls[r * nx];
      for (int x = 0; x + 1 < nx; x++) T[x] += row[x] != row[x + 1];
    }
    std::vector<int> cut;
    for (int x = 0; x + 1 < nx; x++)
      if (8 * T[x] >= (long long)ny * nz) cut.push_back(x + 1);
    if ((int)cut.size() > scn::kMaxSeg - 1) {
      std::stable_sort(cut.begin(), cut.end(), [&](int a, int b) { return T[a - 1] > T[b - 1]; });
      cut.resize(scn::kMaxSeg - 1);
      std::sort(cut.begin(), cut.end());
    }
    seg_x0.insert(seg_x0.end(), cut.begin(), cut.end());
  }
  seg_x0.push_back(nx);
  const int nseg = (int)seg_x0.size() - 1;
  std::vector<long long> seg_base(nseg + 1, 0);
  for (int sg = 0; sg < nseg; sg++)
    seg_base[sg + 1] = (seg_base[sg] + (long long)(seg_x0[sg + 1] - seg_x0[sg]) * ny * nz + 63) / 64 * 64;
  const long long S = seg_base[nseg];  // storage cells, padding included
  if (S > (long long)scn::kCellMask) return fail(h, TLM_ERR_UNSUPPORTED, "more than 2^28 cells per handle");
  std::vector<int> seg_of_x(nx);
  for (int sg = 0; sg < nseg; sg++)
    for (int x = seg_x0[sg]; x < seg_x0[sg + 1]; x++) seg_of_x[x] = sg;
  auto sidx = [&](long long c) -> int32_t {  // reference cell index -> storage index
    const long long r = c / nx;
    const int x = (int)(c - r * nx), sg = seg_of_x[x];
    return (int32_t)(seg_base[sg] + (x - seg_x0[sg]) + (long long)(seg_x0[sg + 1] - seg_x0[sg]) * r);
  };
  // storage class per storage cell (5: padding), class lists in storage order
  std::vector<uint8_t> scls(S, 5);
  for (long long c = 0; c < N; c++) scls[sidx(c)] = cls[c];
  {
    std::vector<uint8_t>().swap(cls);
    auto by_cell = [](const auto& a, const auto& b) { return a.c < b.c; };
    for (auto& r : pr) r.c = sidx(r.c);
    for (auto& r : mr) r.c = sidx(r.c);
    for (auto& r : gr) r.c = sidx(r.c);
    for (auto& r : pmr) r.c = sidx(r.c);
    std::sort(pr.begin(), pr.end(), by_cell);
    std::sort(mr.begin(), mr.end(), by_cell);
    std::sort(gr.begin(), gr.end(), by_cell);
    std::sort(pmr.begin(), pmr.end(), by_cell);
  }

  // ---- 128-byte lines ------------------------------------------------------
  // A line is 128 bytes of one link plane (32 FP32 or 16 FP64 cells).  A
  // line of one class only is written whole by that class's kernel; a mixed
  // line is written whole by the mixed-line kernel, which updates each of
  // its cells with that cell's class code.  No line is ever written in part.
  // Padding cells belong to no line class; a line of padding only is 3.  A
  // line whose cells are free-space cells and material cells (class 2) of
  // one material number only is a material line (4), written by the
  // material-line kernel from chunk records, so that neither its material
  // cells nor its free-space cells go through list words.
  const int L = (int)(128 / link_size(h));
  const long long n_lines = S / L;
  std::vector<uint32_t> line_info(n_lines, 0);
  std::vector<uint8_t> line_pure(n_lines, 0);
  std::vector<uint32_t> mask((S + 31) / 32, 0u);
  // Material lines take materials whose filters are of order <= 1 (PEC,
  // PMC, non-dispersive dielectrics and magnetics); higher-order dispersive
  // materials stay with the mixed-line and material-list kernels, which
  // overlap the line and PML kernels better for them (measured on the
  // 384x384x192 two-pole workload: 1.019 against 0.982 ms per step).
  auto line_material = [&](int m) {
    const tlm_volume_material& vm = h->vmat[m - 1];
    return vm.type != TLM_VMAT_DISPERSIVE ||
           ((!vm.eps_filter_exists || vm.eps_order <= 1) && (!vm.mu_filter_exists || vm.mu_order <= 1));
  };
  std::vector<uint16_t> smat(S, 0);  // material number of the class-2 cells of line materials
  for (const auto& r : mr)
    if (line_material(r.m)) smat[r.c] = r.m;
  long long n_mixed_cells = 0;
  for (long long ln = 0; ln < n_lines; ln++) {
    const long long c0 = ln * L, c1 = c0 + L;
    int k0 = -1, m0 = 0;
    bool same = true, matline = true;
    for (long long c = c0; c < c1; c++) {
      if (scls[c] == 5) continue;
      if (k0 < 0) k0 = scls[c];
      else if (scls[c] != k0) same = false;
      if (scls[c] == 2) {
        if (!smat[c]) matline = false;
        else if (!m0) m0 = smat[c];
        else if (smat[c] != m0) matline = false;
      } else if (scls[c] != 0) {
        matline = false;
      }
    }
    if (k0 < 0) { line_info[ln] = 3; continue; }
    if (same && k0 == 0) { line_info[ln] = 0; continue; }
    for (long long c = c0; c < c1; c++) mask[c >> 5] |= 1u << (c & 31);
    if (matline && m0) { line_info[ln] = 4; continue; }
    if (same) { line_info[ln] = 1; line_pure[ln] = 1; continue; }
    line_info[ln] = 2;
  }
  // material cells of material lines leave the material list
  std::vector<MatRec> mlr;  // storage order
  {
    std::vector<MatRec> keep;
    for (const auto& r : mr) (line_info[r.c / L] == 4 ? mlr : keep).push_back(r);
    mr.swap(keep);
  }
  // Filter state slots (3 per dispersive cell, from 1): first the material
  // lines' cells, chunk by chunk and, inside a chunk, line by line for each
  // material number in the order the material-line records take them, so
  // that the material cells of one record hold consecutive slots; then the
  // material and generic lists in storage order.
  {
    auto disp = [&](int m) { return m > 0 && h->vmat[m - 1].type == TLM_VMAT_DISPERSIVE; };
    long long next = 1;
    std::vector<int32_t> pos_of(mlr.size());
    {
      // material-line records: per 64-cell chunk, materials in order of first
      // appearance over the chunk's material lines
      std::vector<int32_t> order;  // indices into mlr
      order.reserve(mlr.size());
      size_t i = 0;
      while (i < mlr.size()) {
        const long long k = mlr[i].c / scn::kChunk;
        size_t j = i;
        while (j < mlr.size() && mlr[j].c / scn::kChunk == k) j++;
        std::vector<int> ms;
        for (size_t q = i; q < j; q++)
          if (std::find(ms.begin(), ms.end(), (int)mlr[q].m) == ms.end()) ms.push_back(mlr[q].m);
        for (int m : ms)
          for (size_t q = i; q < j; q++)
            if (mlr[q].m == m) order.push_back((int32_t)q);
        i = j;
      }
      for (int32_t q : order)
        if (disp(mlr[q].m)) { mlr[q].f = (int32_t)next; next += 3; }
    }
    std::vector<std::pair<int32_t, int32_t*>> fs;
    for (auto& r : mr) if (disp(r.m)) fs.push_back({r.c, &r.f});
    for (auto& r : gr) if (r.code > 0 && disp(r.m)) fs.push_back({r.c, &r.f});
    std::sort(fs.begin(), fs.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
    for (auto& e : fs) { *e.second = (int32_t)next; next += 3; }
    max_filt = next;
  }
  auto pure_of = [&](int32_t c) { return line_pure[c / L] != 0; };
  std::stable_partition(pr.begin(), pr.end(), [&](const PmlRec& r) { return pure_of(r.c); });
  std::stable_partition(mr.begin(), mr.end(), [&](const MatRec& r) { return pure_of(r.c); });
  std::stable_partition(gr.begin(), gr.end(), [&](const GenRec& r) { return pure_of(r.c); });
  int n_pure_pml = 0, n_pure_mat = 0, n_pure_gen = 0;
  while (n_pure_pml < (int)pr.size() && pure_of(pr[n_pure_pml].c)) n_pure_pml++;
  while (n_pure_mat < (int)mr.size() && pure_of(mr[n_pure_mat].c)) n_pure_mat++;
  while (n_pure_gen < (int)gr.size() && pure_of(gr[n_pure_gen].c)) n_pure_gen++;
  // pure generic cells with code 0 (free-space cells beside a special face)
  // first: the generic kernel's variant without the Thevenin node runs them
  std::stable_partition(gr.begin(), gr.begin() + n_pure_gen, [](const GenRec& r) { return r.code == 0; });
  int n_free_gen = 0;
  while (n_free_gen < n_pure_gen && gr[n_free_gen].code == 0) n_free_gen++;
  // Mixed lines: workgroups of 256 lanes, each owning whole lines.  A
  // special cell takes one lane, bulk cells kMixedBulkPerLane per lane; a
  // group takes consecutive mixed lines (storage order) while its lanes fit.
  // Lanes: special cells sorted by class (PML, generic, material; storage order
  // inside a class), then bulk lanes, bulk lane j holding bulk cells
  // j, j + nb, j + 2 nb, ... of the group (nb = bulk lanes), so that
  // neighbouring lanes touch neighbouring cells.
  constexpr int KB = scn::kMixedBulkPerLane;
  std::vector<int32_t> lds_lines, lds_fd, lds_slot_p, lds_slot_i;
  std::vector<int32_t> slot_a;
  std::vector<std::vector<int32_t>> slot_b(KB > 1 ? KB - 1 : 1);
  {
    int ip = n_pure_pml, im = n_pure_mat, ig = n_pure_gen;
    // class-list positions of the mixed-line cells follow storage order (the
    // tails of the three lists)
    std::vector<long long> mlines;
    std::vector<int32_t> line_first_idx;  // per mixed line: offset into cell_idx
    std::vector<int32_t> cell_idx;
    for (long long ln = 0; ln < n_lines; ln++) {
      if (line_info[ln] != 2) continue;
      mlines.push_back(ln);
      line_first_idx.push_back((int32_t)cell_idx.size());
      const long long c0 = ln * L, c1 = c0 + L;
      for (long long c = c0; c < c1; c++) {
        int k2 = scls[c], idx = 0;
        if (k2 == 1) idx = ip++;
        else if (k2 == 2) idx = im++;
        else if (k2 == 3) idx = ig++;
        cell_idx.push_back(idx);
      }
    }
    if (ip != (int)pr.size() || im != (int)mr.size() || ig != (int)gr.size())
      return fail(h, TLM_ERR_ARG, "internal: mixed-line cells inconsistent");
    static const int rank_of[4] = {3, 0, 2, 1};  // bulk, PML, material, generic
    auto emit = [&](size_t l0, size_t l1) {
      std::vector<std::pair<int, std::pair<int32_t, int32_t>>> spec;  // (rank, (word, idx))
      std::vector<int32_t> bulk;
      for (size_t q = l0; q < l1; q++) {
        const long long c0 = mlines[q] * L, c1 = c0 + L;
        for (long long c = c0; c < c1; c++) {
          const int k2 = scls[c];
          if (k2 >= 4) continue;  // own kernel, or padding
          n_mixed_cells++;
          if (k2 == 0) bulk.push_back((int32_t)c);
          else
            spec.push_back({rank_of[k2], {(int32_t)(c | ((long long)k2 << 28)),
                                          cell_idx[line_first_idx[q] + (c - c0)]}});
        }
      }
      std::stable_sort(spec.begin(), spec.end(), [](const auto& x, const auto& y) { return x.first < y.first; });
      const int nb = (int)((bulk.size() + KB - 1) / KB);
      int lane = 0;
      for (auto& e : spec) {
        slot_a.push_back(e.second.first);
        slot_b[0].push_back(e.second.second);
        for (int k = 1; k < KB - 1; k++) slot_b[k].push_back(-1);
        lane++;
      }
      for (int j = 0; j < nb; j++) {
        slot_a.push_back(bulk[j]);
        for (int k = 1; k < KB; k++) {
          const size_t q = (size_t)j + (size_t)k * nb;
          slot_b[k - 1].push_back(q < bulk.size() ? bulk[q] : -1);
        }
        if (KB == 1) slot_b[0].push_back(-1);
        lane++;
      }
      for (; lane < 256; lane++) {
        slot_a.push_back(-1);
        for (auto& sb : slot_b) sb.push_back(-1);
      }
    };
    size_t l0 = 0;
    int n_spec = 0, n_bulk = 0;
    for (size_t q = 0; q < mlines.size(); q++) {
      const long long c0 = mlines[q] * L, c1 = c0 + L;
      int s = 0, b = 0;
      for (long long c = c0; c < c1; c++)
        if (scls[c] < 4) (scls[c] ? s : b)++;
      if (q > l0 && n_spec + s + (n_bulk + b + KB - 1) / KB > 256) {
        emit(l0, q);
        l0 = q;
        n_spec = 0;
        n_bulk = 0;
      }
      n_spec += s;
      n_bulk += b;
    }
    if (l0 < mlines.size()) emit(l0, mlines.size());
    // LDS-staged workgroups (k_mixed_lds): exactly 256/L consecutive mixed
    // lines each; per cell position its face descriptor (-2: not this
    // kernel's), and a class-sorted slot list of (class << 16 | position).
    constexpr int LC = scn::kLdsCells;
    const int GL = LC / L;
    for (size_t g0 = 0; g0 < mlines.size(); g0 += GL) {
      std::vector<std::pair<int, std::pair<int32_t, int32_t>>> sl;  // (rank, (cls<<16|p, idx))
      for (int q = 0; q < GL; q++) {
        const size_t qi = g0 + q;
        const long long ln = qi < mlines.size() ? mlines[qi] : -1;
        lds_lines.push_back((int32_t)ln);
        for (int k = 0; k < L; k++) {
          const int p = q * L + k;
          const long long c = ln * L + k;
          if (ln < 0 || scls[c] >= 4) {
            lds_fd.push_back(-2);
            continue;
          }
          const int k2 = scls[c];
          const int32_t idx = cell_idx[line_first_idx[qi] + k];
          lds_fd.push_back(k2 == 1 ? pr[idx].fdc : (k2 == 3 ? gr[idx].fdc : -1));
          sl.push_back({rank_of[k2], {(int32_t)((k2 << 16) | p), k2 ? idx : 0}});
        }
      }
      std::stable_sort(sl.begin(), sl.end(), [](const auto& x, const auto& y) { return x.first < y.first; });
      for (int s2 = 0; s2 < LC; s2++) {
        lds_slot_p.push_back(s2 < (int)sl.size() ? sl[s2].second.first : -1);
        lds_slot_i.push_back(s2 < (int)sl.size() ? sl[s2].second.second : 0);
      }
    }
    // Order the workgroups by the kernel variant their classes need
    // (scn::kMixedMasks: without PML cells; with PML cells).
    {
      const size_t ng = lds_lines.size() / GL;
      std::vector<int> var(ng);
      for (size_t g = 0; g < ng; g++) {
        int m = 0;
        for (int s2 = 0; s2 < LC; s2++) {
          const int32_t sp = lds_slot_p[g * LC + s2];
          if (sp >= 0 && (sp >> 16) > 0) m |= 1 << ((sp >> 16) - 1);
        }
        var[g] = (m & 1) ? 1 : 0;
      }
      std::vector<int32_t> l2, f2, p2, i2;
      for (int v = 0; v < 2; v++) {
        h->n_lds_var[v] = 0;
        for (size_t g = 0; g < ng; g++) {
          if (var[g] != v) continue;
          h->n_lds_var[v]++;
          l2.insert(l2.end(), lds_lines.begin() + g * GL, lds_lines.begin() + (g + 1) * GL);
          f2.insert(f2.end(), lds_fd.begin() + g * LC, lds_fd.begin() + (g + 1) * LC);
          p2.insert(p2.end(), lds_slot_p.begin() + g * LC, lds_slot_p.begin() + (g + 1) * LC);
          i2.insert(i2.end(), lds_slot_i.begin() + g * LC, lds_slot_i.begin() + (g + 1) * LC);
        }
      }
      lds_lines.swap(l2);
      lds_fd.swap(f2);
      lds_slot_p.swap(p2);
      lds_slot_i.swap(i2);
    }
  }
  // flat per-class arrays in list order
  std::vector<int32_t> pml_cell, pml_fd;
  std::vector<uint16_t> pml_par;
  std::vector<int2> pml_out_tab, pml_fd_tab;
  for (size_t i = 0; i < pr.size(); i++) {
    int32_t cw = pr[i].c;
    if (pr[i].orow >= 0) { cw |= scn::kFlagOut; pml_out_tab.push_back(make_int2((int)i, pr[i].orow)); }
    if (pr[i].fdc >= 0) { cw |= scn::kFlagFd; pml_fd_tab.push_back(make_int2((int)i, pr[i].fdc)); }
    pml_cell.push_back(cw);
    pml_par.push_back(pr[i].par);
    pml_fd.push_back(pr[i].fdc);
  }
  std::vector<int32_t> mat_cell, mat_filt;
  std::vector<uint16_t> mat_num;
  for (auto& r : mr) { mat_cell.push_back(r.c); mat_num.push_back(r.m); mat_filt.push_back(r.f); }
  std::vector<int32_t> gen_cell, gen_code, gen_mat, gen_filt, gen_exc, gen_out, gen_fd;
  for (auto& r : gr) {
    gen_cell.push_back(r.c); gen_code.push_back(r.code); gen_mat.push_back(r.m); gen_filt.push_back(r.f);
    gen_exc.push_back(r.er); gen_out.push_back(r.orow); gen_fd.push_back(r.fdc);
  }
  h->n_pure_pml = n_pure_pml;
  h->n_pml_fd_pure = h->n_pml_fd_all = 0;
  for (size_t i = 0; i < pr.size(); i++)
    if (pr[i].fdc >= 0) {
      h->n_pml_fd_all++;
      if ((int)i < n_pure_pml) h->n_pml_fd_pure++;
    }
  h->n_pure_mat = n_pure_mat;
  h->n_pure_gen = n_pure_gen;
  h->n_free_gen = n_free_gen;
  h->n_mixed_cells = n_mixed_cells;
  if (exc_rank != h->n_cell_rows)
    return fail(h, TLM_ERR_ARG, "cell excitation rows do not match the cell codes");
  const int n_out_rows = out_rank;

  // ---- excitation CSR ---------------------------------------------------
  std::vector<int32_t> cex_start(h->n_cell_rows + 1, 0), cex_src, fex_start(h->n_face_rows + 1, 0), fex_src;
  std::vector<int32_t> cex_type(6 * (size_t)std::max(h->n_cell_rows, 1), 1), fex_type(6 * (size_t)std::max(h->n_face_rows, 1), 1);
  {
    std::vector<std::vector<int32_t>> cl(h->n_cell_rows), fl(h->n_face_rows);
    for (size_t i = 0; i < h->ps.size(); i++) {
      const tlm_point_source& p = h->ps[i];
      if (p.at_face) {
        fl[p.row - 1].push_back((int32_t)i);
        if (p.component <= 6) fex_type[6 * (p.row - 1) + p.component - 1] = p.source_type;
      } else {
        cl[p.row - 1].push_back((int32_t)i);
        if (p.component <= 6) cex_type[6 * (p.row - 1) + p.component - 1] = p.source_type;
      }
    }
    for (int r = 0; r < h->n_cell_rows; r++) {
      cex_start[r + 1] = cex_start[r] + (int)cl[r].size();
      cex_src.insert(cex_src.end(), cl[r].begin(), cl[r].end());
    }
    for (int r = 0; r < h->n_face_rows; r++) {
      fex_start[r + 1] = fex_start[r] + (int)fl[r].size();
      fex_src.insert(fex_src.end(), fl[r].begin(), fl[r].end());
    }
  }
  std::vector<int32_t> fhy_start(h->n_face_rows + 1, 0), fhy_face;
  {
    std::vector<std::vector<int32_t>> hl(h->n_face_rows);
    for (size_t i = 0; i < h->huy.size(); i++)
      if (h->huy[i].row > 0) hl[h->huy[i].row - 1].push_back((int32_t)i);
    for (int r = 0; r < h->n_face_rows; r++) {
      fhy_start[r + 1] = fhy_start[r] + (int)hl[r].size();
      fhy_face.insert(fhy_face.end(), hl[r].begin(), hl[r].end());
    }
  }
  // Faces excited by exactly one Huygens face with an axis-aligned unit
  // normal, no point source and a free-space (or no) surface material take
  // the flattened descriptor code (value read from the per-step table).
  int n_huy_coded = 0;
  for (auto& code : fd) {
    if (code <= 0) continue;
    const int mt = sf_mtype[code], row = sf_exc[code];
    if (mt != 0 && mt != TLM_SMAT_FREE_SPACE) continue;
    if (sf_out[code] >= 0) continue;  // face outputs read the code's tables
    if (row < 0 || fex_start[row + 1] != fex_start[row] || fhy_start[row + 1] - fhy_start[row] != 1) continue;
    const int f = fhy_face[fhy_start[row]];
    const tlm_huygens_face& hf = h->huy[f];
    const int nz_count = (hf.nx != 0) + (hf.ny != 0) + (hf.nz != 0);
    const int comp = hf.nx ? hf.nx : (hf.ny ? hf.ny : hf.nz);
    if (nz_count != 1 || (comp != 1 && comp != -1) || f >= (1 << scn::kFdHuyFaceBits)) continue;
    const int axis = hf.nx ? 0 : (hf.ny ? 1 : 2);
    code = scn::kFdHuy | (f << 3) | (2 * axis + (comp < 0 ? 1 : 0));
    n_huy_coded++;
  }
  h->n_huy_coded = n_huy_coded;

  // ---- device allocations -------------------------------------------------
  Dev& d = h->d;
  h->P = S + TLM_PLANE_PAD;
  d.nx = nx; d.ny = ny; d.nz = nz; d.N = N; d.P = h->P;
  d.nseg = nseg;
  for (int sg = 0; sg <= nseg; sg++) {
    d.seg_x0[sg] = seg_x0[sg];
    d.seg_base[sg] = seg_base[sg];
  }
  d.dl = h->g.dl; d.dt = h->g.dt; d.c_red = scn::kC0 / h->g.reduced_c_factor;
  d.n_timesteps = h->g.n_timesteps;
  d.R[0] = h->g.r_xmin; d.R[1] = h->g.r_xmax; d.R[2] = h->g.r_ymin;
  d.R[3] = h->g.r_ymax; d.R[4] = h->g.r_zmin; d.R[5] = h->g.r_zmax;
  const size_t ls = link_bytes(h);
  h->n_bufs = pml_cell.empty() ? 2 : 3;
  for (int b = 0; b < h->n_bufs; b++) {
    if ((rc = dev_alloc(h, 12 * (size_t)h->P * ls, &h->planes[b]))) return rc;
    HIP_TRY(h, hipMemsetAsync(h->planes[b], 0, 12 * (size_t)h->P * ls, h->stream));
  }
  if ((rc = dev_upload(h, mask, &d.mask))) return rc;
  h->n_pml = d.n_pml = (int)pml_cell.size();
  if ((rc = dev_upload(h, pml_cell, &d.pml_cell))) return rc;
  if ((rc = dev_upload(h, pml_par, &d.pml_par))) return rc;
  if ((rc = dev_upload(h, pml_out_tab, &d.pml_out_tab))) return rc;
  d.n_pml_out = (int)pml_out_tab.size();
  if ((rc = dev_upload(h, pml_fd_tab, &d.pml_fd_tab))) return rc;
  d.n_pml_fd = (int)pml_fd_tab.size();
  {
    // plain cell indices and descriptors for the connect-only pass
    std::vector<int32_t> cells(pml_cell.size());
    for (size_t i = 0; i < cells.size(); i++) cells[i] = pml_cell[i] & scn::kCellMask;
    if ((rc = dev_upload(h, cells, &h->d_pml_cell))) return rc;
    if ((rc = dev_upload(h, pml_fd, &h->d_pml_fd))) return rc;
  }
  d.line_len = L;
  if ((rc = dev_upload(h, line_info, &d.line_info))) return rc;
  {
    // k_bulk_chunks records: 64-cell chunks holding a pure bulk line
    std::vector<int4> ch;
    const int lpc = scn::kChunk / L;
    for (long long k = 0; k < (S + scn::kChunk - 1) / scn::kChunk; k++) {
      int m = 0;
      for (int q = 0; q < lpc; q++)
        if (k * lpc + q < n_lines && line_info[k * lpc + q] == 0) m |= 1 << q;
      if (!m) continue;
      const long long c0 = k * scn::kChunk;
      int sg = 0;
      while (sg + 1 < nseg && c0 >= seg_base[sg + 1]) sg++;
      const long long w = seg_x0[sg + 1] - seg_x0[sg];
      const long long r = c0 - seg_base[sg], row = r / w;
      ch.push_back(make_int4((int)(k << 4) | m, (int)(seg_x0[sg] + (r - row * w)) | (sg << 16), (int)(row % ny),
                             (int)(row / ny)));
    }
    if (k_max_chunks_exceeded(S)) return fail(h, TLM_ERR_UNSUPPORTED, "storage too large for the chunk table");
    h->n_bulk_chunks = (int)ch.size();
    if ((rc = dev_upload(h, ch, &h->d_bulk_chunks))) return rc;
    // k_pml_chunks records (two int4 each): per 64-cell chunk holding a pure
    // PML line, its pure PML lines, the chunk's first-cell coordinates, the
    // list position of the first of their cells and whether any of them has
    // an output row or a special face.  The pure part of the PML list is in
    // storage order, so the cells of a chunk's lines hold consecutive
    // positions.
    std::vector<int32_t> line_pos(n_lines, -1);
    std::vector<uint8_t> line_flag(n_lines, 0);
    for (int i = 0; i < n_pure_pml; i++) {
      const long long ln = pr[i].c / L;
      if (line_pos[ln] < 0) line_pos[ln] = i;
      if (pr[i].orow >= 0 || pr[i].fdc >= 0) line_flag[ln] = 1;
    }
    std::vector<int4> pc, pcf;  // unflagged, flagged
    long long covered = 0;
    for (long long k = 0; k < (S + scn::kChunk - 1) / scn::kChunk; k++) {
      int m = 0, first = -1, flag = 0, expect = -1;
      for (int q = 0; q < lpc && k * lpc + q < n_lines; q++) {
        const long long ln = k * lpc + q;
        if (line_pos[ln] < 0) continue;
        if (first < 0) first = expect = line_pos[ln];
        if (line_pos[ln] != expect) return fail(h, TLM_ERR_ARG, "PML list positions of a chunk are not consecutive");
        long long nc = 0;
        for (long long c = ln * L; c < (ln + 1) * L; c++) nc += scls[c] != 5;
        expect += (int)nc;
        covered += nc;
        m |= 1 << q;
        flag |= line_flag[ln];
      }
      if (!m) continue;
      const long long c0 = k * scn::kChunk;
      int sg = 0;
      while (sg + 1 < nseg && c0 >= seg_base[sg + 1]) sg++;
      const long long w = seg_x0[sg + 1] - seg_x0[sg];
      const long long r = c0 - seg_base[sg], row = r / w;
      auto& to = flag ? pcf : pc;
      to.push_back(make_int4((int)(k << 4) | m, (int)(seg_x0[sg] + (r - row * w)) | (sg << 16), (int)(row % ny),
                             (int)(row / ny)));
      to.push_back(make_int4(first, flag, 0, 0));
    }
    if (covered != n_pure_pml) return fail(h, TLM_ERR_ARG, "PML chunk records do not cover the pure PML lines");
    h->n_pml_chunks_fast = (int)(pc.size() / 2);
    pc.insert(pc.end(), pcf.begin(), pcf.end());
    h->n_pml_chunks = (int)(pc.size() / 2);
    if ((rc = dev_upload(h, pc, &h->d_pml_chunks))) return rc;
    // k_mat_lines records (two int4 each): per 64-cell chunk and material
    // number, its material lines, the chunk's first-cell coordinates, the
    // material, the lanes that are material cells and their first filter
    // slot
    std::vector<int4> mrec;
    std::vector<int32_t> slot_of(0);
    size_t im = 0;  // walks mlr (storage order)
    h->n_matline_cells = h->n_matline_mat = 0;
    h->matline_filter_bytes = 0;
    for (long long k = 0; k < (S + scn::kChunk - 1) / scn::kChunk; k++) {
      const long long c0 = k * scn::kChunk;
      std::vector<std::pair<int, int>> mats;  // (material, line mask)
      for (int q = 0; q < lpc && k * lpc + q < n_lines; q++) {
        const long long ln = k * lpc + q;
        if (line_info[ln] != 4) continue;
        int m = 0;
        for (long long c = ln * L; c < (ln + 1) * L && !m; c++)
          if (scls[c] == 2) m = smat[c];
        size_t j = 0;
        while (j < mats.size() && mats[j].first != m) j++;
        if (j == mats.size()) mats.push_back({m, 0});
        mats[j].second |= 1 << q;
      }
      if (mats.empty()) continue;
      int sg = 0;
      while (sg + 1 < nseg && c0 >= seg_base[sg + 1]) sg++;
      const long long w = seg_x0[sg + 1] - seg_x0[sg];
      const long long r = c0 - seg_base[sg], row = r / w;
      const int4 head0 = make_int4(0, (int)(seg_x0[sg] + (r - row * w)) | (sg << 16), (int)(row % ny),
                                   (int)(row / ny));
      for (const auto& mq : mats) {
        uint64_t lanes = 0;
        int32_t slot0 = 0;
        bool first = true;
        for (int q = 0; q < lpc; q++) {
          if (!((mq.second >> q) & 1)) continue;
          for (int t = q * L; t < (q + 1) * L; t++) {
            const long long c = c0 + t;
            if (scls[c] == 5) continue;
            h->n_matline_cells++;
            if (scls[c] != 2) continue;
            while (im < mlr.size() && mlr[im].c < c) im++;
            if (im >= mlr.size() || mlr[im].c != c) return fail(h, TLM_ERR_ARG, "internal: material line cells");
            lanes |= 1ull << t;
            if (first) { slot0 = mlr[im].f; first = false; }
            h->n_matline_mat++;
            h->matline_filter_bytes += filter_bytes_of(h, mq.first);
          }
        }
        int4 r0 = head0;
        r0.x = (int)(k << 4) | mq.second;
        mrec.push_back(r0);
        mrec.push_back(make_int4(mq.first, (int)(uint32_t)lanes, (int)(uint32_t)(lanes >> 32), slot0));
      }
    }
    // records of materials whose filters have order <= 1 first (the
    // kernel variant without the general node update), the others after
    {
      auto is_low = [&](int m) {
        const tlm_volume_material& vm = h->vmat[m - 1];
        return vm.type == TLM_VMAT_DISPERSIVE && (!vm.eps_filter_exists || vm.eps_order <= 1) &&
               (!vm.mu_filter_exists || vm.mu_order <= 1);
      };
      std::vector<int4> lo, hi;
      for (size_t k = 0; k < mrec.size(); k += 2) {
        auto& to = is_low(mrec[k + 1].x) ? lo : hi;
        to.push_back(mrec[k]);
        to.push_back(mrec[k + 1]);
      }
      h->n_mat_recs_low = (int)(lo.size() / 2);
      lo.insert(lo.end(), hi.begin(), hi.end());
      mrec.swap(lo);
    }
    h->n_mat_recs = (int)(mrec.size() / 2);
    if ((rc = dev_upload(h, mrec, &h->d_mat_recs))) return rc;
  }
  {
    // Workgroups of the bulk line kernel: the 256-cell storage blocks that
    // hold a pure bulk line.  The y-band table splits the blocks of every
    // z-plane of a segment into eight equal runs in storage order and gives
    // run k to XCD k (workgroup b % 8 == k), so a block's y-1 and z-1
    // neighbour rows were read just before by the same XCD (its own L2) and
    // every XCD gets the same share of each plane.
    const int lpb = 256 / L;
    h->bulk_blocks.clear();
    std::vector<long long> plane_of;  // (segment, z) key per listed block
    for (long long b = 0; b < (S + 255) / 256; b++) {
      long long first = -1;
      for (int q = 0; q < lpb && first < 0 && b * lpb + q < n_lines; q++)
        if (line_info[b * lpb + q] == 0) first = (b * lpb + q) * L;
      if (first < 0) continue;
      h->bulk_blocks.push_back((int32_t)b);
      int sg = 0;
      while (sg + 1 < nseg && first >= seg_base[sg + 1]) sg++;
      const long long row = (first - seg_base[sg]) / (seg_x0[sg + 1] - seg_x0[sg]);
      plane_of.push_back((long long)sg * nz + row / ny);
    }
    std::vector<std::vector<int32_t>> band(8);
    for (size_t i0 = 0; i0 < plane_of.size();) {
      size_t i1 = i0;
      while (i1 < plane_of.size() && plane_of[i1] == plane_of[i0]) i1++;
      const size_t m = i1 - i0;
      for (size_t j = 0; j < m; j++) band[j * 8 / m].push_back(h->bulk_blocks[i0 + j]);
      i0 = i1;
    }
    size_t per = 0;
    for (auto& v : band) per = std::max(per, v.size());
    std::vector<int32_t> yb(8 * per, -1);
    for (int k = 0; k < 8; k++)
      for (size_t j = 0; j < band[k].size(); j++) yb[8 * j + k] = band[k][j];
    if ((rc = dev_upload(h, h->bulk_blocks, &h->d_bulk_blocks))) return rc;
    if ((rc = dev_upload(h, yb, &h->d_bulk_blocks_yband))) return rc;
    h->n_bulk_blocks_yband = (int)yb.size();
  }
  if ((rc = dev_upload(h, lds_lines, &h->d_lds_lines))) return rc;
  if ((rc = dev_upload(h, lds_fd, &h->d_lds_fd))) return rc;
  if ((rc = dev_upload(h, lds_slot_p, &h->d_lds_slot_p))) return rc;
  if ((rc = dev_upload(h, lds_slot_i, &h->d_lds_slot_i))) return rc;
  h->n_lds_wg = (int)(lds_slot_p.size() / scn::kLdsCells);
  {
    std::vector<int32_t> sb;
    for (auto& v : slot_b) sb.insert(sb.end(), v.begin(), v.end());
    if ((rc = dev_upload(h, slot_a, &h->d_slot_cell))) return rc;
    if ((rc = dev_upload(h, sb, &h->d_slot_idx))) return rc;
    h->n_mixed_blocks = (int)(slot_a.size() / 256);
  }
  if ((rc = dev_upload(h, h->pml_par, &d.pml_params))) return rc;
  // state planes start on 128-byte lines: a multiple of 64 list positions
  d.n_pml_state = (std::max(d.n_pml, 1) + 63) / 64 * 64;
  {
    void* p;
    const bool plain = h->form == TLM_FORM_PLAIN_FP64;
    const size_t words = plain ? scn::PS_N : scn::PC_N;
    if ((rc = dev_alloc(h, words * d.n_pml_state * sizeof(double), &p))) return rc;
    HIP_TRY(h, hipMemsetAsync(p, 0, words * d.n_pml_state * sizeof(double), h->stream));
    d.pml_state = (double*)p;
    d.pml_hi = nullptr;
    d.pml_lo = nullptr;
    if (!plain) {
      const size_t nw = (size_t)scn::SW_N * d.n_pml_state;
      if ((rc = dev_alloc(h, nw * 6, &p))) return rc;
      HIP_TRY(h, hipMemsetAsync(p, 0, nw * 6, h->stream));
      d.pml_hi = (uint32_t*)p;
      d.pml_lo = (uint16_t*)((char*)p + nw * 4);
    }
    if ((rc = dev_alloc(h, (size_t)3 * d.n_pml_state * sizeof(double), &p))) return rc;
    HIP_TRY(h, hipMemsetAsync(p, 0, (size_t)3 * d.n_pml_state * sizeof(double), h->stream));
    d.pml_It = (double*)p;
  }
  h->n_mat = d.n_mat = (int)mat_cell.size();
  if ((rc = dev_upload(h, mat_cell, &d.mat_cell))) return rc;
  if ((rc = dev_upload(h, mat_num, &d.mat_num))) return rc;
  if ((rc = dev_upload(h, mat_filt, &d.mat_filt))) return rc;
  {
    std::vector<int32_t> pc, pp, pm, po, pf;
    for (auto& r : pmr) { pc.push_back(r.c); pp.push_back(r.par); pm.push_back(r.m); po.push_back(r.orow); pf.push_back(r.fdc); }
    h->n_pm = d.n_pm = (int)pmr.size();
    if ((rc = dev_upload(h, pc, &d.pm_cell))) return rc;
    if ((rc = dev_upload(h, pp, &d.pm_par))) return rc;
    if ((rc = dev_upload(h, pm, &d.pm_mat))) return rc;
    if ((rc = dev_upload(h, po, &d.pm_out))) return rc;
    if ((rc = dev_upload(h, pf, &d.pm_fd))) return rc;
    if (h->n_pm) {
      void* p;
      const size_t bytes = (size_t)scn::PM_N * h->n_pm * sizeof(double);
      if ((rc = dev_alloc(h, bytes, &p))) return rc;
      HIP_TRY(h, hipMemsetAsync(p, 0, bytes, h->stream));
      d.pm_state = (double*)p;
    }
  }
  h->n_gen = d.n_gen = (int)gen_cell.size();
  if ((rc = dev_upload(h, gen_cell, &d.gen_cell))) return rc;
  if ((rc = dev_upload(h, gen_code, &d.gen_code))) return rc;
  if ((rc = dev_upload(h, gen_mat, &d.gen_mat))) return rc;
  if ((rc = dev_upload(h, gen_filt, &d.gen_filt))) return rc;
  if ((rc = dev_upload(h, gen_exc, &d.gen_exc))) return rc;
  if ((rc = dev_upload(h, gen_out, &d.gen_out))) return rc;
  if ((rc = dev_upload(h, gen_fd, &d.gen_fd))) return rc;
  h->d_gen_cell = d.gen_cell;
  {
    // host-path cells: storage indices and 1-based coordinates, scan order
    std::vector<int32_t> hs;
    h->host_xyz.clear();
    for (int32_t c : host_ref) {
      hs.push_back(sidx(c));
      const long long r = c / nx;
      h->host_xyz.push_back((int32_t)(c - r * nx) + 1);
      h->host_xyz.push_back((int32_t)(r % ny) + 1);
      h->host_xyz.push_back((int32_t)(r / ny) + 1);
    }
    h->n_host = (int)hs.size();
    if ((rc = dev_upload(h, hs, &h->d_host_cells))) return rc;
    if (h->n_host) {
      void* p;
      if ((rc = dev_alloc(h, (size_t)12 * h->n_host * sizeof(double), &p))) return rc;
      h->d_host_buf = (double*)p;
    }
  }
  h->d_gen_fd = d.gen_fd;
  if ((rc = dev_upload(h, h->vmat, &d.vmat))) return rc;
  d.n_filt = std::max<long long>(max_filt, 1);
  {
    void* p;
    size_t fb = (size_t)scn::NW * d.n_filt * sizeof(double);
    if ((rc = dev_alloc(h, fb, &p))) return rc;
    HIP_TRY(h, hipMemsetAsync(p, 0, fb, h->stream));
    d.eps_w = (double*)p;
    if ((rc = dev_alloc(h, fb, &p))) return rc;
    HIP_TRY(h, hipMemsetAsync(p, 0, fb, h->stream));
    d.mu_w = (double*)p;
  }
  if ((rc = dev_upload(h, fd, &d.fd))) return rc;
  if ((rc = dev_upload(h, sf_mtype, &d.sf_mtype))) return rc;
  if ((rc = dev_upload(h, sf_exc, &d.sf_exc))) return rc;
  if ((rc = dev_upload(h, sf_out, &d.sf_out))) return rc;
  if ((rc = dev_upload(h, sf_host, &d.sf_host))) return rc;
  {
    h->n_host_faces = (int)hf_codes.size();
    h->host_face_xyz.clear();
    std::vector<int32_t> cp, cm, ax;
    for (int code : hf_codes) {
      const long long c = sf_cell[code];
      const int a = sf_axis[code];
      const long long off = a == 0 ? 1 : (a == 1 ? nx : (long long)nx * ny);
      cp.push_back(sidx(c));
      cm.push_back(sidx(c - off));
      ax.push_back(a);
      const long long r = c / nx;
      h->host_face_xyz.push_back((int32_t)(c - r * nx) + 1);
      h->host_face_xyz.push_back((int32_t)(r % ny) + 1);
      h->host_face_xyz.push_back((int32_t)(r / ny) + 1);
      h->host_face_xyz.push_back(a + 1);
    }
    if ((rc = dev_upload(h, cp, &h->d_hf_cp))) return rc;
    if ((rc = dev_upload(h, cm, &h->d_hf_cm))) return rc;
    if ((rc = dev_upload(h, ax, &h->d_hf_ax))) return rc;
    void* p = nullptr;
    if ((rc = dev_alloc(h, (size_t)4 * std::max(h->n_host_faces, 1) * sizeof(double), &p))) return rc;
    h->d_hf_tot = (double*)p;
    HIP_TRY(h, hipMemsetAsync(p, 0, (size_t)4 * std::max(h->n_host_faces, 1) * sizeof(double), h->stream));
    if ((rc = dev_alloc(h, (size_t)4 * std::max(h->n_host_faces, 1) * sizeof(double), &p))) return rc;
    h->d_hf_buf = (double*)p;
    d.host_tot = h->d_hf_tot;
  }
  d.n_face_out_rows = n_face_out_rows;
  if (n_face_out_rows) {
    void* p;
    if ((rc = dev_alloc(h, (size_t)h->cap * n_face_out_rows * 12 * sizeof(double), &p))) return rc;
    d.face_ring = (double*)p;
  }
  if ((rc = dev_upload(h, h->value, &d.value))) return rc;
  if ((rc = dev_upload(h, h->value_face, &d.value_face))) return rc;
  if ((rc = dev_upload(h, cex_start, &d.cex_start))) return rc;
  if ((rc = dev_upload(h, cex_src, &d.cex_src))) return rc;
  if ((rc = dev_upload(h, cex_type, &d.cex_type))) return rc;
  if ((rc = dev_upload(h, fex_start, &d.fex_start))) return rc;
  if ((rc = dev_upload(h, fex_src, &d.fex_src))) return rc;
  if ((rc = dev_upload(h, fex_type, &d.fex_type))) return rc;
  if ((rc = dev_upload(h, fhy_start, &d.fhy_start))) return rc;
  if ((rc = dev_upload(h, fhy_face, &d.fhy_face))) return rc;
  if ((rc = dev_upload(h, h->ps, &d.ps))) return rc;
  if ((rc = dev_upload(h, h->huy, &d.hf))) return rc;
  d.n_huy = (int)h->huy.size();
  if (n_huy_coded) {
    void* p;
    if ((rc = dev_alloc(h, h->huy.size() * sizeof(double), &p))) return rc;
    d.huy_val = (double*)p;
  }
  d.huy_function = h->huy_function;
  d.huy_offset_min = h->huy_offset_min;
  for (int k = 0; k < 3; k++) { d.Ei[k] = h->Ei[k]; d.Hi[k] = h->Hi[k]; }
  d.n_out_rows = n_out_rows;
  {
    void* p;
    if ((rc = dev_alloc(h, (size_t)h->cap * std::max(n_out_rows, 1) * 6 * sizeof(double), &p))) return rc;
    d.ring = (double*)p;
  }
  // probes and DFT
  for (size_t i = 0; i < h->pt_row.size(); i++)
    if (h->pt_row[i] < 1 || h->pt_row[i] > n_out_rows || h->pt_comp[i] < 1 || h->pt_comp[i] > 6)
      return fail(h, TLM_ERR_ARG, "output point row/component out of range");
  {
    std::vector<int32_t> r0(h->pt_row.size()), c0(h->pt_comp.size());
    for (size_t i = 0; i < r0.size(); i++) { r0[i] = h->pt_row[i] - 1; c0[i] = h->pt_comp[i] - 1; }
    const int32_t* pr;
    if ((rc = dev_upload(h, r0, &pr))) return rc;
    h->d_prow = (int32_t*)pr;
    if ((rc = dev_upload(h, c0, &pr))) return rc;
    h->d_pcomp = (int32_t*)pr;
    const double* fq;
    if ((rc = dev_upload(h, h->freq, &fq))) return rc;
    h->d_freq = (double*)fq;
    void* p;
    size_t nb = std::max<size_t>(1, h->pt_row.size() * h->freq.size()) * 2 * sizeof(double);
    if ((rc = dev_alloc(h, nb, &p))) return rc;
    HIP_TRY(h, hipMemsetAsync(p, 0, nb, h->stream));
    h->d_dft = (double*)p;
  }
  if (!h->fo.empty() || !h->ff.empty()) {
    // frequency outputs: items refer to this handle's cell and face output rows
    std::vector<scn::FreqOut> outs;
    std::vector<scn::FreqItem> items;
    size_t k = 0;
    for (size_t o = 0; o < h->fo.size(); o++) {
      const tlm_frequency_output& f = h->fo[o];
      outs.push_back({f.frequency, 1.0, f.kind, f.field_component, f.first_timestep, f.last_timestep,
                      f.timestep_interval});
      const int nrows = f.kind == 1 ? d.n_out_rows : d.n_face_out_rows;
      for (int i = 0; i < f.n_items; i++, k++) {
        const int r = h->fo_rows[k];
        if (r < 1 || r > nrows)
          return fail(h, TLM_ERR_ARG, "frequency output " + std::to_string(o + 1) + ": output row out of range");
        const int side = f.kind == 2 ? h->fo_sides[k] : 1, nrm = f.kind == 2 ? h->fo_normals[k] : 0;
        if (f.kind == 2 && (side < 1 || side > 2 || nrm == 0 || nrm < -3 || nrm > 3))
          return fail(h, TLM_ERR_ARG, "frequency output " + std::to_string(o + 1) + ": bad face side or normal");
        items.push_back({(int)o, r, side, nrm});
        (f.kind == 1 ? h->n_fo_vol : h->n_fo_surf)++;
      }
    }
    // far-field surfaces: J (components 7..9) and M (10..12) of every face
    // at every step, times dt (face_output_far_field, far_field_surfaces.F90:
    // 252-351)
    h->ff_item0 = (int)items.size();
    size_t kf = 0, most = 1;
    h->d_ff_centres.clear();
    for (size_t sfc = 0; sfc < h->ff.size(); sfc++) {
      const tlm_far_field_surface& f = h->ff[sfc];
      for (int c = 7; c <= 12; c++) {
        const int o = (int)outs.size();
        outs.push_back({f.frequency, h->g.dt, 2, c, 1, 2147483647, 1});
        for (int i = 0; i < f.n_faces; i++) {
          const int r = h->ff_rows[kf + i];
          if (r < 1 || r > d.n_face_out_rows)
            return fail(h, TLM_ERR_ARG, "far-field surface " + std::to_string(sfc + 1) + ": output row out of range");
          items.push_back({o, r, h->ff_sides[kf + i], h->ff_normals[kf + i]});
          h->n_fo_surf++;
        }
      }
      std::vector<double> cen(h->ff_centres.begin() + 3 * kf, h->ff_centres.begin() + 3 * (kf + f.n_faces));
      const double* dc;
      if ((rc = dev_upload(h, cen, &dc))) return rc;
      h->d_ff_centres.push_back(dc);
      most = std::max(most, (size_t)f.n_theta * f.n_phi);
      kf += f.n_faces;
    }
    if (!h->ff.empty()) {
      void* pe;
      if ((rc = dev_alloc(h, most * 4 * sizeof(double), &pe))) return rc;
      h->d_ff_e = (double*)pe;
    }
    h->n_fo_items = (int)items.size();
    if ((rc = dev_upload(h, outs, &h->d_fo))) return rc;
    if ((rc = dev_upload(h, items, &h->d_fo_items))) return rc;
    void* p;
    const size_t nb = (size_t)std::max(1, h->n_fo_items) * 2 * sizeof(double);
    if ((rc = dev_alloc(h, nb, &p))) return rc;
    HIP_TRY(h, hipMemsetAsync(p, 0, nb, h->stream));
    h->d_fo_acc = (double*)p;
  }
  if (!h->fd.empty()) {
    // time-domain field outputs: per output its items and a ring of at
    // most the output capacity of frames (fewer when its window has fewer)
    size_t k = 0;
    for (size_t o = 0; o < h->fd.size(); o++) {
      const tlm_field_output& f = h->fd[o];
      const int nrows = f.kind == 1 ? d.n_out_rows : d.n_face_out_rows;
      std::vector<scn::FreqItem> items;
      for (int i = 0; i < f.n_items; i++, k++) {
        const int r = h->fd_rows[k];
        if (r < 1 || r > nrows)
          return fail(h, TLM_ERR_ARG, "field output " + std::to_string(o + 1) + ": output row out of range");
        items.push_back({(int)o, r, f.kind == 2 ? h->fd_sides[k] : 1, f.kind == 2 ? h->fd_normals[k] : 0});
      }
      const scn::FreqItem* di;
      if ((rc = dev_upload(h, items, &di))) return rc;
      int due = 0;
      for (int t = std::max(1, f.first_timestep); t <= std::min(h->g.n_timesteps, f.last_timestep) && due < h->cap; t++)
        if ((t - f.first_timestep) % f.timestep_interval == 0) due++;
      const int cap = std::max(1, due);
      void* p;
      if ((rc = dev_alloc(h, (size_t)cap * std::max(1, f.n_items) * sizeof(double), &p))) return rc;
      h->d_fd_items.push_back(di);
      h->d_fd_ring.push_back((double*)p);
      h->fd_cap.push_back(cap);
      (f.kind == 1 ? h->n_fd_vol : h->n_fd_surf)++;
    }
    h->fd_used.assign(h->fd.size(), 0);
    h->fd_first.assign(h->fd.size(), 0);
    h->fd_host_frames.assign(h->fd.size(), 0);
    h->fd_host.assign(h->fd.size(), {});
  }
  if (h->rcs_on) {
    // faces sorted by d = K.r (initialise_rcs_surfaces, RCS_surfaces.F90:
    // 159-367): K the unit vector of (theta, phi), r the face centre
    const double kx = sin(h->rcs.theta) * cos(h->rcs.phi), ky = sin(h->rcs.theta) * sin(h->rcs.phi),
                 kz = cos(h->rcs.theta);
    h->rcs_faces.clear();
    for (int i = 0; i < h->rcs.n_faces; i++) {
      const int r = h->rcs_rows[i];
      if (r < 1 || r > d.n_face_out_rows) return fail(h, TLM_ERR_ARG, "RCS surface: output row out of range");
      const double* c = &h->rcs_centres[3 * (size_t)i];
      h->rcs_faces.push_back({r, h->rcs_sides[i], h->rcs_normals[i], 0, kx * c[0] + ky * c[1] + kz * c[2]});
    }
    std::stable_sort(h->rcs_faces.begin(), h->rcs_faces.end(),
                     [](const scn::RcsFace& a, const scn::RcsFace& b) { return a.d < b.d; });
    if ((rc = dev_upload(h, h->rcs_faces, &h->d_rcs_faces))) return rc;
    void* p;
    const size_t nb = (size_t)h->rcs.n_points * 4 * sizeof(double);
    if ((rc = dev_alloc(h, nb, &p))) return rc;
    HIP_TRY(h, hipMemsetAsync(p, 0, nb, h->stream));
    h->d_rcs = (double*)p;
  }
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  h->finalized = true;
  return TLM_OK;
}

// Move field output o's frames from its device ring to the host.
static int drain_field_output(tlm_handle* h, size_t o) {
  const int n = h->fd_used[o];
  if (!n) return TLM_OK;
  const size_t items = (size_t)h->fd[o].n_items;
  std::vector<double>& v = h->fd_host[o];
  const size_t old = v.size();
  v.resize(old + (size_t)n * items);
  if (items) {
    HIP_TRY(h, hipMemcpyAsync(v.data() + old, h->d_fd_ring[o], (size_t)n * items * sizeof(double),
                              hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(h, hipStreamSynchronize(h->stream));
  }
  h->fd_host_frames[o] += n;
  h->fd_used[o] = 0;
  return TLM_OK;
}

// One frame of every field output of this kind due at `timestep`
// (get_output_flag), from cell-output slot or face-output block `slot`.
int launch_field_outputs(tlm_handle* h, int kind, int slot, int timestep) {
  if (!(kind == 1 ? h->n_fd_vol : h->n_fd_surf) || slot < 0) return TLM_OK;
  const double* src = kind == 1 ? h->d.ring + (size_t)slot * h->d.n_out_rows * 6
                                : h->d.face_ring + (size_t)slot * h->d.n_face_out_rows * 12;
  for (size_t o = 0; o < h->fd.size(); o++) {
    const tlm_field_output& f = h->fd[o];
    if (f.kind != kind || timestep < f.first_timestep || timestep > f.last_timestep ||
        (timestep - f.first_timestep) % f.timestep_interval != 0)
      continue;
    int rc;
    if (h->fd_used[o] == h->fd_cap[o] && (rc = drain_field_output(h, o))) return rc;
    if (h->fd_used[o] == 0 && h->fd_host_frames[o] == 0) h->fd_first[o] = timestep;
    if (f.n_items)
      hipLaunchKernelGGL(scn::k_field_out, dim3((f.n_items + 255) / 256), dim3(256), 0, h->stream, h->d_fd_items[o],
                         f.n_items, kind, f.field_component, f.both_sides, src,
                         h->d_fd_ring[o] + (size_t)h->fd_used[o] * f.n_items);
    h->fd_used[o]++;
  }
  return TLM_OK;
}

// The RCS surface's contributions of the face outputs in block `slot` of
// the connect of step `timestep` (face_output_rcs, RCS_surfaces.F90:386-559).
void launch_rcs(tlm_handle* h, int slot, int timestep) {
  if (!h->rcs_on || slot < 0 || h->rcs_faces.empty()) return;
  const double dt = h->g.dt, dl = h->g.dl;
  const double time = (double)(timestep - 1) * dt + dt / 2.0;
  scn::RcsArgs a;
  a.ct = cos(h->rcs.theta);
  a.st = sin(h->rcs.theta);
  a.cp = cos(h->rcs.phi);
  a.sp = sin(h->rcs.phi);
  a.dA = dl * dl;
  a.time_c0 = time * scn::kC0;
  a.dl = dl;
  a.points_per_cell = h->rcs.points_per_cell;
  a.offset = h->rcs.point_offset;
  a.nf = (int)h->rcs_faces.size();
  a.rows = h->d.n_face_out_rows;
  // points of the faces with the smallest and largest K.r
  const int p_hi = scn::rcs_point(h->rcs_faces.front().d, a.time_c0, a.points_per_cell, dl, a.offset);
  const int p_lo = scn::rcs_point(h->rcs_faces.back().d, a.time_c0, a.points_per_cell, dl, a.offset);
  const int lo = std::max(p_lo, 1), hi = std::min(p_hi, h->rcs.n_points);
  if (lo > hi) return;
  a.p_lo = lo;
  hipLaunchKernelGGL(scn::k_rcs, dim3(hi - lo + 1), dim3(256), 0, h->stream, a, h->d_rcs_faces,
                     h->d.face_ring + (size_t)slot * h->d.n_face_out_rows * 12, h->d_rcs);
}

// Frequency outputs of one kind (1: volumes over the cell-output rows of
// ring slot `slot` at step `timestep`; 2: surfaces over the face-output
// block `slot` of the connect of step `timestep`).
void launch_freq_outputs(tlm_handle* h, int kind, int slot, int timestep) {
  if (!(kind == 1 ? h->n_fo_vol : h->n_fo_surf) || slot < 0) return;
  const double dt = h->g.dt;
  const double* src;
  int rows;
  double time;
  if (kind == 1) {
    rows = h->d.n_out_rows;
    src = h->d.ring + (size_t)slot * rows * 6;
    time = (double)(timestep - 1) * dt;  // cell_output (run_TLM.F90:240)
  } else {
    rows = h->d.n_face_out_rows;
    src = h->d.face_ring + (size_t)slot * rows * 12;
    time = (double)(timestep - 1) * dt + dt / 2.0;  // face_output (run_TLM.F90:248)
  }
  hipLaunchKernelGGL(scn::k_freq_out, dim3((h->n_fo_items + 255) / 256), dim3(256), 0, h->stream, h->d_fo,
                     h->d_fo_items, h->n_fo_items, kind, src, rows, timestep, time, h->d_fo_acc);
}

// Move filled ring slots to the host staging vector.
int drain_ring(tlm_handle* h) {
  if (h->ring_used == 0) return TLM_OK;
  const int rows = h->d.n_out_rows;
  size_t n = (size_t)h->ring_used * rows * 6;
  size_t old = h->host_out.size();
  if (h->host_steps == 0) h->host_first = h->ring_first_step;
  h->host_out.resize(old + n);
  HIP_TRY(h, hipMemcpyAsync(h->host_out.data() + old, h->d.ring, n * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  h->host_steps += h->ring_used;
  h->ring_used = 0;
  return TLM_OK;
}

// Move filled face-output slots to the host staging vector.
int drain_face_ring(tlm_handle* h) {
  if (h->fring_used == 0) return TLM_OK;
  const size_t w = (size_t)h->d.n_face_out_rows * 12;
  const size_t n = (size_t)h->fring_used * w;
  const size_t old = h->host_fout.size();
  if (h->host_fsteps == 0) h->host_ffirst = h->fring_first_step;
  h->host_fout.resize(old + n);
  HIP_TRY(h, hipMemcpyAsync(h->host_fout.data() + old, h->d.face_ring, n * sizeof(double), hipMemcpyDeviceToHost,
                            h->stream));
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  h->host_fsteps += h->fring_used;
  h->fring_used = 0;
  return TLM_OK;
}

// Face-output slot for the connect of `timestep` (-1 when no face outputs).
int reserve_face_slot(tlm_handle* h, int timestep, int* slot) {
  *slot = -1;
  if (h->d.n_face_out_rows == 0) return TLM_OK;
  int rc;
  if (h->fring_used == h->cap && (rc = drain_face_ring(h))) return rc;
  if (h->fring_used == 0) h->fring_first_step = timestep;
  *slot = h->fring_used++;
  return TLM_OK;
}

// Bulk schedule (tlm_set_schedule): variant 3, the line kernel, unless
// chosen otherwise per handle; variant 0: 64 x rows tiles of one plane;
// variant 1: z-marching blocks of zc planes; variant 2: 16-byte vector
// accesses.
bool ext_form(const tlm_handle* h) { return h->form == TLM_FORM_EXTENDED_FP32; }

int sched_variant(const tlm_handle* h) {
  if (h->stz_n > 0) return 3;  // the line kernel applies the small-value flush
  if (h->form == TLM_FORM_EXTENDED_FP32) return 3;  // the only bulk kernel built for it
  if (h->sched_variant >= 0) return h->sched_variant;
  return 3;
}

int bulk_vec_width(const tlm_handle* h) {
  int w = h->form == TLM_FORM_PLAIN_FP64 ? 2 : 4;
  if (sched_variant(h) != 2 || h->nx % w != 0 || h->P % w != 0) return 1;
  return w;
}

scn::BulkMap bulk_map(const tlm_handle* h) {
  scn::BulkMap m;
  m.by = h->sched_by;
  const int w = bulk_vec_width(h);
  m.nbx = (h->nx + 64 * w - 1) / (64 * w);
  m.nby = (h->ny + m.by - 1) / m.by;
  long long tiles = (long long)m.nbx * m.nby;
  if (sched_variant(h) != 1) {
    m.zc = 1;
  } else {
    int zc = h->sched_zc;
    while (zc > 2 && tiles * ((h->nz + zc - 1) / zc) < 4096) zc >>= 1;
    m.zc = std::max(1, std::min(zc, h->nz));
  }
  m.nchunks = (h->nz + m.zc - 1) / m.zc;
  m.nblocks = (int)(tiles * m.nchunks);
  m.xcd_remap = (h->sched_xcd && m.nblocks % 8 == 0) ? 1 : 0;
  return m;
}

// Per-step Huygens face values for the kFdHuy descriptor codes, at the
// pending connect's timestep (before any kernel that pulls across faces).
void launch_huygens_values(tlm_handle* h, const PullArgs& a) {
  if (!h->n_huy_coded) return;
  hipLaunchKernelGGL(scn::k_huygens_values, dim3((h->d.n_huy + 255) / 256), dim3(256), 0, h->stream, h->d,
                     a.ts_connect);
}

template <typename T, bool PLAIN>
int launch_scatter(tlm_handle* h, int timestep, bool pull, const PullArgs& a, int slot) {
  const Dev& d = h->d;
  if (pull) launch_huygens_values(h, a);
  const T* Lin = (const T*)h->planes[h->cur];
  T* Lout = (T*)h->planes[out_buffer(h)];
  // The material, generic and PML-material kernels write cells disjoint
  // from the bulk, PML and mixed-line kernels' and read only the input
  // buffer, so with two streams they run on the second stream concurrently
  // with those (joined before the next kernel of the handle's stream).
  // Profiling serialises them.  The bulk kernel goes first (measured
  // faster).
  hipStream_t cs = h->stream;
  const bool fork = h->streams > 1 && !h->profiling;
  if (fork) {
    if (!h->side) {
      HIP_TRY(h, hipStreamCreateWithFlags(&h->side, hipStreamNonBlocking));
      HIP_TRY(h, hipEventCreateWithFlags(&h->ev_fork, hipEventDisableTiming));
      HIP_TRY(h, hipEventCreateWithFlags(&h->ev_join, hipEventDisableTiming));
    }
    HIP_TRY(h, hipEventRecord(h->ev_fork, h->stream));
    HIP_TRY(h, hipStreamWaitEvent(h->side, h->ev_fork, 0));
    cs = h->side;
  }
  {
    ProfScope ps(h, TLM_KCLASS_BULK);
    constexpr bool ext = std::is_same<T, scn::xlink>::value;
    if (ext || sched_variant(h) == 3) {
      // one storage block per workgroup over the whole storage (measured
      // faster than a block table), or the block tables; the extended form
      // has the chunk kernel only (tlm_set_schedule refuses the others)
      const bool yband = !ext && h->sched_xcd == 2, direct = ext || (!yband && h->sched_by == 1);
      const int32_t* blocks = direct ? nullptr : (yband ? h->d_bulk_blocks_yband : h->d_bulk_blocks);
      const int nb = yband ? h->n_bulk_blocks_yband : (int)h->bulk_blocks.size();
      const int per = h->sched_by;  // storage blocks per workgroup
      const int nwg = direct ? h->n_bulk_chunks : (yband ? 8 * ((nb / 8 + per - 1) / per) : (nb + per - 1) / per);
      if (direct) {
        if (!nwg) {
        } else if (pull)
          hipLaunchKernelGGL((scn::k_bulk_chunks<T, PLAIN, true>), dim3(nwg), dim3(scn::kChunk), 0, h->stream, d, Lin,
                             Lout, a, h->d_bulk_chunks, h->sched_xcd);
        else
          hipLaunchKernelGGL((scn::k_bulk_chunks<T, PLAIN, false>), dim3(nwg), dim3(scn::kChunk), 0, h->stream, d, Lin,
                             Lout, a, h->d_bulk_chunks, h->sched_xcd);
      } else if constexpr (ext) {
      } else if (nb) {
        if (pull)
          hipLaunchKernelGGL((scn::k_bulk_blocks<T, PLAIN, true>), dim3(nwg), dim3(256), 0, h->stream, d, Lin, Lout, a,
                             blocks, nb, per, h->sched_xcd);
        else
          hipLaunchKernelGGL((scn::k_bulk_blocks<T, PLAIN, false>), dim3(nwg), dim3(256), 0, h->stream, d, Lin, Lout, a,
                             blocks, nb, per, h->sched_xcd);
      }
    } else if constexpr (!ext) {
      scn::BulkMap m = bulk_map(h);
      dim3 bb(64, m.by, 1);
      const int w = bulk_vec_width(h);
      if (w > 1) {
        constexpr int W = sizeof(T) == 4 ? 4 : 2;
        if (pull) hipLaunchKernelGGL((scn::k_bulk_vec<T, PLAIN, true, W>), dim3(m.nblocks), bb, 0, h->stream, d, Lin, Lout, a, m);
        else hipLaunchKernelGGL((scn::k_bulk_vec<T, PLAIN, false, W>), dim3(m.nblocks), bb, 0, h->stream, d, Lin, Lout, a, m);
      } else if (m.zc == 1) {
        if (pull) hipLaunchKernelGGL((scn::k_bulk_cell<T, PLAIN, true>), dim3(m.nblocks), bb, 0, h->stream, d, Lin, Lout, a, m);
        else hipLaunchKernelGGL((scn::k_bulk_cell<T, PLAIN, false>), dim3(m.nblocks), bb, 0, h->stream, d, Lin, Lout, a, m);
      } else {
        if (pull) hipLaunchKernelGGL((scn::k_bulk_z<T, PLAIN, true>), dim3(m.nblocks), bb, 0, h->stream, d, Lin, Lout, a, m);
        else hipLaunchKernelGGL((scn::k_bulk_z<T, PLAIN, false>), dim3(m.nblocks), bb, 0, h->stream, d, Lin, Lout, a, m);
      }
    }
  }
  if (h->n_mat_recs) {
    // order <= 1 materials' records, then the others
    ProfScope ps(h, TLM_KCLASS_MATERIAL);
    const int nl = h->n_mat_recs_low, nh = h->n_mat_recs - nl;
    if (nl && pull)
      hipLaunchKernelGGL((scn::k_mat_lines<T, PLAIN, true, true>), dim3(nl), dim3(scn::kChunk), 0, cs, d, Lin, Lout, a,
                         h->d_mat_recs);
    else if (nl)
      hipLaunchKernelGGL((scn::k_mat_lines<T, PLAIN, false, true>), dim3(nl), dim3(scn::kChunk), 0, cs, d, Lin, Lout,
                         a, h->d_mat_recs);
    if (nh && pull)
      hipLaunchKernelGGL((scn::k_mat_lines<T, PLAIN, true, false>), dim3(nh), dim3(scn::kChunk), 0, cs, d, Lin, Lout,
                         a, h->d_mat_recs + 2 * nl);
    else if (nh)
      hipLaunchKernelGGL((scn::k_mat_lines<T, PLAIN, false, false>), dim3(nh), dim3(scn::kChunk), 0, cs, d, Lin, Lout,
                         a, h->d_mat_recs + 2 * nl);
  }
  if (h->n_pure_mat) {
    ProfScope ps(h, TLM_KCLASS_MATERIAL);
    dim3 g((h->n_pure_mat + 255) / 256);
    if (pull) hipLaunchKernelGGL((scn::k_material<T, PLAIN, true>), g, dim3(256), 0, cs, d, Lin, Lout, a, h->n_pure_mat);
    else hipLaunchKernelGGL((scn::k_material<T, PLAIN, false>), g, dim3(256), 0, cs, d, Lin, Lout, a, h->n_pure_mat);
  }
  if (h->n_pure_gen) {
    // code-0 cells (free space beside a special face) on the variant without
    // the Thevenin node, then the others
    ProfScope ps(h, TLM_KCLASS_GENERIC);
    const int nf = h->n_free_gen, nr = h->n_pure_gen - nf;
    if (nf && pull)
      hipLaunchKernelGGL((scn::k_generic<T, PLAIN, true, true>), dim3((nf + 255) / 256), dim3(256), 0, cs, d, Lin, Lout,
                         a, timestep, slot, 0, nf);
    else if (nf)
      hipLaunchKernelGGL((scn::k_generic<T, PLAIN, false, true>), dim3((nf + 255) / 256), dim3(256), 0, cs, d, Lin,
                         Lout, a, timestep, slot, 0, nf);
    if (nr && pull)
      hipLaunchKernelGGL((scn::k_generic<T, PLAIN, true, false>), dim3((nr + 255) / 256), dim3(256), 0, cs, d, Lin,
                         Lout, a, timestep, slot, nf, h->n_pure_gen);
    else if (nr)
      hipLaunchKernelGGL((scn::k_generic<T, PLAIN, false, false>), dim3((nr + 255) / 256), dim3(256), 0, cs, d, Lin,
                         Lout, a, timestep, slot, nf, h->n_pure_gen);
  }
  if (h->n_pm) {
    ProfScope ps(h, TLM_KCLASS_PML_MATERIAL);
    dim3 g((h->n_pm + 255) / 256);
    if (pull) hipLaunchKernelGGL((scn::k_pml_material<T, PLAIN, true>), g, dim3(256), 0, cs, d, Lin, Lout, a, slot);
    else hipLaunchKernelGGL((scn::k_pml_material<T, PLAIN, false>), g, dim3(256), 0, cs, d, Lin, Lout, a, slot);
  }
  if (h->n_pml_chunks) {
    // unflagged chunks on the fast kernel, flagged ones on the general one
    ProfScope ps(h, TLM_KCLASS_PML);
    const int nf = h->n_pml_chunks_fast;
    if (nf && pull)
      hipLaunchKernelGGL((scn::k_pml_chunks<T, PLAIN, true, true>), dim3(nf), dim3(scn::kChunk), 0, h->stream, d, Lin,
                         Lout, a, slot, h->d_pml_chunks);
    else if (nf)
      hipLaunchKernelGGL((scn::k_pml_chunks<T, PLAIN, false, true>), dim3(nf), dim3(scn::kChunk), 0, h->stream, d, Lin,
                         Lout, a, slot, h->d_pml_chunks);
    const int ng = h->n_pml_chunks - nf;
    if (ng && pull)
      hipLaunchKernelGGL((scn::k_pml_chunks<T, PLAIN, true, false>), dim3(ng), dim3(scn::kChunk), 0, h->stream, d, Lin,
                         Lout, a, slot, h->d_pml_chunks + 2 * nf);
    else if (ng)
      hipLaunchKernelGGL((scn::k_pml_chunks<T, PLAIN, false, false>), dim3(ng), dim3(scn::kChunk), 0, h->stream, d,
                         Lin, Lout, a, slot, h->d_pml_chunks + 2 * nf);
  }
  // Mixed lines on the main stream after the PML kernels: their 256-lane
  // workgroups of up to 148 registers find no room beside the bulk kernel
  // and wait for it to drain, holding up the side stream's material and
  // generic kernels (measured 0.462 -> 0.453 ms per step on the 256^3
  // workload against the side stream).
  const int mixed = ext_form(h) ? 1 : h->mixed_variant >= 0 ? h->mixed_variant : 1;  // LDS-staged: measured faster
  if (mixed == 1 && h->n_lds_wg) {
    ProfScope ps(h, TLM_KCLASS_MIXED);
    int wg0 = 0;
    for (int v = 0; v < 2; v++) {
      const int n = h->n_lds_var[v];
      if (!n) continue;
      auto go = [&](auto kern) {
        hipLaunchKernelGGL(kern, dim3(n), dim3(scn::kLdsCells), 0, h->stream, d, Lin, Lout, a, h->d_lds_lines, h->d_lds_fd,
                           h->d_lds_slot_p, h->d_lds_slot_i, timestep, slot, wg0);
      };
      if (pull) {
        if (v == 0) go(scn::k_mixed_lds<T, PLAIN, true, scn::kMixedMasks[0]>);
        else go(scn::k_mixed_lds<T, PLAIN, true, scn::kMixedMasks[1]>);
      } else {
        if (v == 0) go(scn::k_mixed_lds<T, PLAIN, false, scn::kMixedMasks[0]>);
        else go(scn::k_mixed_lds<T, PLAIN, false, scn::kMixedMasks[1]>);
      }
      wg0 += n;
    }
  } else if (mixed == 0 && h->n_mixed_blocks) {
    ProfScope ps(h, TLM_KCLASS_MIXED);
    dim3 g(h->n_mixed_blocks);
    if constexpr (std::is_same<T, scn::xlink>::value) {
    } else if (pull)
      hipLaunchKernelGGL((scn::k_mixed_lines<T, PLAIN, true>), g, dim3(256), 0, h->stream, d, Lin, Lout, a,
                         h->d_slot_cell, h->d_slot_idx, (long long)h->n_mixed_blocks * 256, timestep, slot);
    else
      hipLaunchKernelGGL((scn::k_mixed_lines<T, PLAIN, false>), g, dim3(256), 0, h->stream, d, Lin, Lout, a,
                         h->d_slot_cell, h->d_slot_idx, (long long)h->n_mixed_blocks * 256, timestep, slot);
  }
  if (fork) {
    HIP_TRY(h, hipEventRecord(h->ev_join, h->side));
    HIP_TRY(h, hipStreamWaitEvent(h->stream, h->ev_join, 0));
  }
  return TLM_OK;
}

template <typename T, bool PLAIN>
int launch_connect(tlm_handle* h, const PullArgs& a) {
  const Dev& d = h->d;
  launch_huygens_values(h, a);
  const T* Lin = (const T*)h->planes[h->cur];
  T* Lout = (T*)h->planes[out_buffer(h)];
  dim3 bb(64, 4, 1);
  dim3 gb((d.nx + 63) / 64, (d.ny + 3) / 4, d.nz);
  hipLaunchKernelGGL((scn::k_connect_bulk<T, PLAIN>), gb, bb, 0, h->stream, d, Lin, Lout, a);
  if (h->n_pml)
    hipLaunchKernelGGL((scn::k_connect_list<T, PLAIN>), dim3((h->n_pml + 255) / 256), dim3(256), 0, h->stream, d, Lin,
                       Lout, a, h->d_pml_cell, h->d_pml_fd, h->n_pml);
  if (h->n_gen)
    hipLaunchKernelGGL((scn::k_connect_list<T, PLAIN>), dim3((h->n_gen + 255) / 256), dim3(256), 0, h->stream, d, Lin,
                       Lout, a, h->d_gen_cell, h->d_gen_fd, h->n_gen);
  if (h->n_mat)
    hipLaunchKernelGGL((scn::k_connect_list<T, PLAIN>), dim3((h->n_mat + 255) / 256), dim3(256), 0, h->stream, d, Lin,
                       Lout, a, d.mat_cell, nullptr, h->n_mat);
  if (h->n_pm)
    hipLaunchKernelGGL((scn::k_connect_list<T, PLAIN>), dim3((h->n_pm + 255) / 256), dim3(256), 0, h->stream, d, Lin,
                       Lout, a, d.pm_cell, d.pm_fd, h->n_pm);
  return TLM_OK;
}

// One scatter at `timestep`, consuming any pending connect.
int do_scatter(tlm_handle* h, int timestep) {
  int rc;
  if ((rc = finalize(h))) return rc;
  if (timestep < 1 || timestep > h->g.n_timesteps) return fail(h, TLM_ERR_ARG, "timestep out of range");
  bool pull;
  if (h->kind == Kind::INCIDENT) {
    pull = false;
  } else {
    if (!h->pend_connect) return fail(h, TLM_ERR_STATE, "scatter called twice without connect");
    pull = true;
  }
  const double small = (h->stz_n > 0 && timestep % h->stz_n == 0) ? h->stz_value : 0.0;
  int fslot = -1;
  if (pull && (rc = reserve_face_slot(h, h->pend_ts, &fslot))) return rc;
  PullArgs a{h->pend_boundary, h->pend_ts, small, fslot};
  a.prev = h->planes[h->prv];
  a.prev_mode = h->prv_mode;
  a.prev_apply_r = h->prv_apply_r;
  a.prev_small = h->prv_small;
  if (h->d.n_out_rows > 0 && h->ring_used == h->cap)
    if ((rc = drain_ring(h))) return rc;
  int slot = h->ring_used;
  rc = with_form(h->form, [&](auto t, auto plain) {
    return launch_scatter<decltype(t), decltype(plain)::value>(h, timestep, pull, a, slot);
  });
  if (rc) return rc;
  HIP_TRY(h, hipGetLastError());
  if (h->d.n_out_rows > 0) {
    if (h->ring_used == 0) h->ring_first_step = timestep;
    h->ring_used++;
    int np = (int)h->pt_row.size(), nf = (int)h->freq.size();
    if (np && nf) {
      ProfScope ps(h, TLM_KCLASS_DFT);
      hipLaunchKernelGGL(scn::k_dft, dim3((np * nf + 255) / 256), dim3(256), 0, h->stream, h->d.ring, h->d.n_out_rows,
                         slot, timestep, h->d_prow, h->d_pcomp, np, h->d_freq, nf, h->g.dt, h->d_dft);
    }
    launch_freq_outputs(h, 1, slot, timestep);
    if ((rc = launch_field_outputs(h, 1, slot, timestep))) return rc;
  }
  // the pulled connect's face outputs
  if (pull) {
    launch_freq_outputs(h, 2, fslot, a.ts_connect);
    launch_rcs(h, fslot, a.ts_connect);
    if ((rc = launch_field_outputs(h, 2, fslot, a.ts_connect))) return rc;
  }
  const int out = out_buffer(h);
  h->prv = h->cur;
  h->cur = out;
  h->prv_mode = pull ? 1 : 2;
  h->prv_apply_r = a.apply_r;
  h->prv_small = small;
  h->kind = Kind::SCATTERED;
  h->pend_connect = false;
  h->pend_boundary = false;
  return TLM_OK;
}

// Make the buffer hold incident voltages (apply any pending connect).
int materialize(tlm_handle* h) {
  int rc;
  if ((rc = finalize(h))) return rc;
  if (h->kind == Kind::INCIDENT) return TLM_OK;
  if (!h->pend_connect) return TLM_OK;  // scattered state is what V holds
  int fslot = -1;
  if ((rc = reserve_face_slot(h, h->pend_ts, &fslot))) return rc;
  PullArgs a{h->pend_boundary, h->pend_ts, 0.0, fslot};
  rc = with_form(h->form, [&](auto t, auto plain) {
    return launch_connect<decltype(t), decltype(plain)::value>(h, a);
  });
  if (rc) return rc;
  launch_freq_outputs(h, 2, fslot, a.ts_connect);
  launch_rcs(h, fslot, a.ts_connect);
  if ((rc = launch_field_outputs(h, 2, fslot, a.ts_connect))) return rc;
  HIP_TRY(h, hipGetLastError());
  h->cur = out_buffer(h);  // prv stays: the next step still pulls it again
  h->kind = Kind::INCIDENT;
  h->pend_connect = false;
  h->pend_boundary = false;
  return TLM_OK;
}

}  // namespace engine

// ===========================================================================
// C ABI
// ===========================================================================

// Staging buffer of tlm_download_links_noalloc: a file-scope device array,
// one copy per GPU for the life of the process.  Host code never passes the
// array's name to a kernel or a copy (on the host it only stands for the
// symbol); it uses the per-device address from hipGetSymbolAddress.
constexpr size_t kStagingDoubles = TLM_NOALLOC_STAGING_DOUBLES;
static __device__ __align__(256) double scratch[kStagingDoubles];

namespace {
constexpr int kMaxStagingDevices = 64;
std::mutex g_staging_mu;
double* g_staging[kMaxStagingDevices] = {};

// The staging buffer's device address on the handle's GPU, looked up once
// per device with that device current.
int staging_buffer(tlm_handle* h, double** out) {
  if (h->device < 0 || h->device >= kMaxStagingDevices)
    return fail(h, TLM_ERR_UNSUPPORTED, "device number beyond the staging table");
  std::lock_guard<std::mutex> lock(g_staging_mu);
  if (!g_staging[h->device]) {
    int prev = -1;
    HIP_TRY(h, hipGetDevice(&prev));
    if (prev != h->device) HIP_TRY(h, hipSetDevice(h->device));
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (prev != h->device) (void)hipSetDevice(prev);
    if (e != hipSuccess) return hip_fail(h, e, "hipGetSymbolAddress(staging buffer)");
    g_staging[h->device] = (double*)p;
  }
  *out = g_staging[h->device];
  return TLM_OK;
}
}  // namespace

extern "C" {

int tlm_create(const tlm_grid* g, int32_t form, tlm_handle** out) {
  if (!g || !out) return TLM_ERR_ARG;
  *out = nullptr;
  if (g->nx < 1 || g->ny < 1 || g->nz < 1 || g->n_timesteps < 1 || !(g->dl > 0) || !(g->dt > 0))
    return TLM_ERR_ARG;
  if (!valid_form(form)) return TLM_ERR_ARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return TLM_ERR_HIP;
  tlm_handle* h = new (std::nothrow) tlm_handle();
  if (!h) return TLM_ERR_NOMEM;
  h->g = *g;
  h->form = form;
  h->nx = g->nx; h->ny = g->ny; h->nz = g->nz;
  h->N = (long long)g->nx * g->ny * g->nz;
  h->P = (h->N + 63) / 64 * 64;
  if (hipGetDevice(&h->device) != hipSuccess ||
      hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
    delete h;
    return TLM_ERR_HIP;
  }
  *out = h;
  return TLM_OK;
}

void tlm_destroy(tlm_handle* h) {
  if (!h) return;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  if (h->side) (void)hipStreamSynchronize(h->side);
  for (auto& e : h->ev) {
    (void)hipEventDestroy(e.second.first);
    (void)hipEventDestroy(e.second.second);
  }
  for (auto& b : h->bufs) (void)hipFree(b.p);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  if (h->side) (void)hipStreamDestroy(h->side);
  if (h->ev_fork) (void)hipEventDestroy(h->ev_fork);
  if (h->ev_join) (void)hipEventDestroy(h->ev_join);
  delete h;
}

const char* tlm_last_error(const tlm_handle* h) { return h ? h->err.c_str() : "null handle"; }

int64_t tlm_device_bytes(const tlm_handle* h) { return h ? h->dev_bytes : 0; }

int tlm_set_cell_codes(tlm_handle* h, const int32_t* code, int32_t n_special, const int32_t* to_material,
                       const int32_t* to_pml, const int32_t* to_excitation, const int32_t* to_output,
                       const int32_t* to_cable) {
  if (!h) return TLM_ERR_ARG;
  if (h->finalized) return fail(h, TLM_ERR_STATE, "tables are fixed after the first step");
  if (!code || n_special < 0) return fail(h, TLM_ERR_ARG, "bad cell code table");
  h->cell_code.assign(code, code + h->N);
  h->n_special = n_special;
  auto cp = [&](const int32_t* src, std::vector<int32_t>& dst) {
    if (src) dst.assign(src, src + n_special); else dst.assign(n_special, 0);
  };
  cp(to_material, h->c_mat);
  if (to_material) h->c_filt.assign(to_material + n_special, to_material + 2 * (size_t)n_special);
  else h->c_filt.assign(n_special, 0);
  cp(to_pml, h->c_pml);
  cp(to_excitation, h->c_exc);
  cp(to_output, h->c_out);
  cp(to_cable, h->c_cable);
  h->cells_set = true;
  return TLM_OK;
}

int tlm_set_face_codes(tlm_handle* h, const int32_t* code, int32_t n_special, const int32_t* to_material,
                       const int32_t* to_excitation, const int32_t* to_output, const int32_t* to_cable) {
  if (!h) return TLM_ERR_ARG;
  if (h->finalized) return fail(h, TLM_ERR_STATE, "tables are fixed after the first step");
  long long nf = (long long)(h->nx - 1) * h->ny * h->nz + (long long)h->nx * (h->ny - 1) * h->nz +
                 (long long)h->nx * h->ny * (h->nz - 1);
  if (n_special < 0 || (nf > 0 && !code)) return fail(h, TLM_ERR_ARG, "bad face code table");
  if (code) h->face_code.assign(code, code + nf); else h->face_code.assign(nf, 0);
  h->n_special_faces = n_special;
  auto cp = [&](const int32_t* src, std::vector<int32_t>& dst) {
    if (src) dst.assign(src, src + n_special); else dst.assign(n_special, 0);
  };
  cp(to_material, h->f_mat);
  cp(to_excitation, h->f_exc);
  cp(to_output, h->f_out);
  // A cable face (to_cable != 0) is connected as its surface material or
  // excitation makes it; the reference then calls face_cable_junction
  // (connect_normal_to_x.F90:317-319), which reads and writes only the cable
  // model's own state, so the host runs it after tlm_connect.
  cp(to_cable, h->f_cable);
  h->faces_set = true;
  return TLM_OK;
}

int tlm_set_volume_materials(tlm_handle* h, int32_t n, const tlm_volume_material* m) {
  if (!h || n < 0 || (n > 0 && !m)) return TLM_ERR_ARG;
  if (h->finalized) return fail(h, TLM_ERR_STATE, "tables are fixed after the first step");
  for (int i = 0; i < n; i++) {
    if (m[i].type < 1 || m[i].type > 3) return fail(h, TLM_ERR_ARG, "unknown volume material type");
    if (m[i].eps_order < 0 || m[i].eps_order > TLM_MAX_FILTER_ORDER || m[i].mu_order < 0 ||
        m[i].mu_order > TLM_MAX_FILTER_ORDER)
      return fail(h, TLM_ERR_ARG, "filter order out of range");
    if ((m[i].eps_filter_exists && m[i].eps_b[0] == 0.0) || (m[i].mu_filter_exists && m[i].mu_b[0] == 0.0))
      return fail(h, TLM_ERR_ARG, "filter b0 is zero");
  }
  h->vmat.assign(m, m + n);
  return TLM_OK;
}

int tlm_set_surface_materials(tlm_handle* h, int32_t n, const int32_t* type) {
  if (!h || n < 0 || (n > 0 && !type)) return TLM_ERR_ARG;
  if (h->finalized) return fail(h, TLM_ERR_STATE, "tables are fixed after the first step");
  for (int i = 0; i < n; i++)
    if (type[i] != TLM_SMAT_PEC && type[i] != TLM_SMAT_PMC && type[i] != TLM_SMAT_FREE_SPACE && !host_smat(type[i]))
      return fail(h, TLM_ERR_UNSUPPORTED, "unknown surface material type");
  h->smat.assign(type, type + n);
  return TLM_OK;
}

int tlm_set_pml(tlm_handle* h, int32_t n_params, const tlm_pml_parameter* p, int32_t n_cells,
                const int32_t* cell_param) {
  if (!h || n_params < 0 || n_cells < 0 || (n_params && !p) || (n_cells && !cell_param)) return TLM_ERR_ARG;
  if (h->finalized) return fail(h, TLM_ERR_STATE, "tables are fixed after the first step");
  h->pml_par.assign(p, p + n_params);
  h->pml_cell_param.assign(cell_param, cell_param + n_cells);
  return TLM_OK;
}

int tlm_set_excitation_functions(tlm_handle* h, int32_t n_functions, const double* value, const double* value_face) {
  if (!h || n_functions < 0 || (n_functions && (!value || !value_face))) return TLM_ERR_ARG;
  if (h->finalized) return fail(h, TLM_ERR_STATE, "tables are fixed after the first step");
  size_t n = (size_t)n_functions * h->g.n_timesteps;
  h->n_functions = n_functions;
  h->value.assign(value, value + n);
  h->value_face.assign(value_face, value_face + n);
  return TLM_OK;
}

int tlm_set_point_sources(tlm_handle* h, int32_t n, const tlm_point_source* s, int32_t n_cell_rows,
                          int32_t n_face_rows) {
  if (!h || n < 0 || (n && !s) || n_cell_rows < 0 || n_face_rows < 0) return TLM_ERR_ARG;
  if (h->finalized) return fail(h, TLM_ERR_STATE, "tables are fixed after the first step");
  for (int i = 0; i < n; i++) {
    if (s[i].function < 1 || s[i].function > h->n_functions)
      return fail(h, TLM_ERR_ARG, "point source function out of range");
    if (s[i].component < 1 || s[i].component > (s[i].at_face ? 12 : 9))
      return fail(h, TLM_ERR_ARG, "point source component out of range");
    if (s[i].source_type != 0 && s[i].source_type != 1) return fail(h, TLM_ERR_ARG, "bad source type");
    int rows = s[i].at_face ? n_face_rows : n_cell_rows;
    if (s[i].row < 1 || s[i].row > rows) return fail(h, TLM_ERR_ARG, "point source row out of range");
  }
  h->ps.assign(s, s + n);
  h->n_cell_rows = n_cell_rows;
  h->n_face_rows = n_face_rows;
  return TLM_OK;
}

int tlm_set_huygens_surface(tlm_handle* h, int32_t function, double offset_min, const double* Ei, const double* Hi,
                            int32_t n_faces, const tlm_huygens_face* faces) {
  if (!h || n_faces < 0 || (n_faces && (!faces || !Ei || !Hi))) return TLM_ERR_ARG;
  if (h->finalized) return fail(h, TLM_ERR_STATE, "tables are fixed after the first step");
  if (n_faces && (function < 1 || function > h->n_functions)) return fail(h, TLM_ERR_ARG, "Huygens function out of range");
  for (int i = 0; i < n_faces; i++) {
    if (faces[i].row < 0 || faces[i].row > h->n_face_rows) return fail(h, TLM_ERR_ARG, "Huygens face row out of range");
    int s = abs(faces[i].nx) + abs(faces[i].ny) + abs(faces[i].nz);
    if (s != 1) return fail(h, TLM_ERR_ARG, "Huygens face normal must be a unit axis vector");
  }
  h->huy_function = function;
  h->huy_offset_min = offset_min;
  for (int k = 0; k < 3; k++) { h->Ei[k] = Ei ? Ei[k] : 0.0; h->Hi[k] = Hi ? Hi[k] : 0.0; }
  h->huy.assign(faces, faces + n_faces);
  return TLM_OK;
}

int tlm_set_output_points(tlm_handle* h, int32_t n, const int32_t* row, const int32_t* component) {
  if (!h || n < 0 || (n && (!row || !component))) return TLM_ERR_ARG;
  if (h->finalized) return fail(h, TLM_ERR_STATE, "tables are fixed after the first step");
  h->pt_row.assign(row, row + n);
  h->pt_comp.assign(component, component + n);
  return TLM_OK;
}

int tlm_set_dft_frequencies(tlm_handle* h, int32_t n, const double* f) {
  if (!h || n < 0 || (n && !f)) return TLM_ERR_ARG;
  if (h->finalized) return fail(h, TLM_ERR_STATE, "tables are fixed after the first step");
  h->freq.assign(f, f + n);
  return TLM_OK;
}

int tlm_set_frequency_outputs(tlm_handle* h, int32_t n, const tlm_frequency_output* outs, const int32_t* rows,
                              const int32_t* sides, const int32_t* normals) {
  if (!h || n < 0 || (n && !outs)) return TLM_ERR_ARG;
  if (h->finalized) return fail(h, TLM_ERR_STATE, "tables are fixed after the first step");
  size_t items = 0;
  for (int i = 0; i < n; i++) {
    const tlm_frequency_output& f = outs[i];
    if ((f.kind != 1 && f.kind != 2) || f.n_items < 0 || f.timestep_interval < 1 || f.field_component < 1 ||
        f.field_component > (f.kind == 1 ? 6 : 9))
      return fail(h, TLM_ERR_ARG, "frequency output " + std::to_string(i + 1) + ": bad kind, component or window");
    items += (size_t)f.n_items;
  }
  if (items && (!rows || !sides || !normals)) return TLM_ERR_ARG;
  h->fo.assign(outs, outs + n);
  h->fo_rows.assign(rows, rows + items);
  h->fo_sides.assign(sides, sides + items);
  h->fo_normals.assign(normals, normals + items);
  return TLM_OK;
}

int tlm_read_frequency_output(tlm_handle* h, int32_t k, double* re_im) {
  if (!h || k < 0 || !re_im) return TLM_ERR_ARG;
  int rc;
  if ((rc = finalize(h))) return rc;
  if (k >= (int)h->fo.size()) return fail(h, TLM_ERR_ARG, "no such frequency output");
  if ((rc = materialize(h))) return rc;  // a pending connect adds its face outputs
  size_t off = 0;
  for (int i = 0; i < k; i++) off += (size_t)h->fo[i].n_items;
  const size_t n = (size_t)h->fo[k].n_items * 2;
  if (n) {
    HIP_TRY(h, hipMemcpyAsync(re_im, h->d_fo_acc + off * 2, n * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(h, hipStreamSynchronize(h->stream));
  }
  return TLM_OK;
}

int tlm_set_field_outputs(tlm_handle* h, int32_t n, const tlm_field_output* outs, const int32_t* rows,
                          const int32_t* sides, const int32_t* normals) {
  if (!h || n < 0 || (n && !outs)) return TLM_ERR_ARG;
  if (h->finalized) return fail(h, TLM_ERR_STATE, "tables are fixed after the first step");
  size_t items = 0;
  for (int i = 0; i < n; i++) {
    const tlm_field_output& f = outs[i];
    const int c = f.field_component;
    const bool surf_c = (c >= 1 && c <= 9) || (c >= 13 && c <= 15);
    const bool ok = (f.kind == 1 && c >= 1 && c <= 6 && !f.both_sides) ||
                    (f.kind == 2 && surf_c && (!f.both_sides || (c >= 7 && c <= 9) || c == 15));
    if (!ok || f.n_items < 0 || f.timestep_interval < 1 || (f.both_sides != 0 && f.both_sides != 1))
      return fail(h, TLM_ERR_ARG, "field output " + std::to_string(i + 1) + ": bad kind, component or window");
    items += (size_t)f.n_items;
  }
  if (items && (!rows || !sides || !normals)) return TLM_ERR_ARG;
  for (size_t i = 0, k = 0; i < (size_t)n; i++)
    for (int j = 0; j < outs[i].n_items; j++, k++)
      if (outs[i].kind == 2 && (sides[k] < 1 || sides[k] > 2 || normals[k] == 0 || normals[k] < -3 || normals[k] > 3))
        return fail(h, TLM_ERR_ARG, "field output " + std::to_string(i + 1) + ": bad face side or normal");
  h->fd.assign(outs, outs + n);
  h->fd_rows.assign(rows, rows + items);
  h->fd_sides.assign(sides, sides + items);
  h->fd_normals.assign(normals, normals + items);
  return TLM_OK;
}

int tlm_read_field_output(tlm_handle* h, int32_t k, double* values, int32_t max_frames, int32_t* n_frames,
                          int32_t* first_timestep) {
  if (!h || k < 0 || max_frames < 0 || !n_frames || !first_timestep || (max_frames && !values)) return TLM_ERR_ARG;
  int rc;
  if ((rc = finalize(h))) return rc;
  if (k >= (int)h->fd.size()) return fail(h, TLM_ERR_ARG, "no such field output");
  if ((rc = materialize(h))) return rc;  // a pending connect records its face frame
  if ((rc = drain_field_output(h, (size_t)k))) return rc;
  const tlm_field_output& f = h->fd[k];
  const int n = std::min(max_frames, h->fd_host_frames[k]);
  const size_t items = (size_t)f.n_items;
  std::vector<double>& v = h->fd_host[k];
  if (n && items) std::memcpy(values, v.data(), (size_t)n * items * sizeof(double));
  *n_frames = n;
  *first_timestep = h->fd_first[k];
  v.erase(v.begin(), v.begin() + (size_t)n * items);
  h->fd_host_frames[k] -= n;
  h->fd_first[k] += n * f.timestep_interval;
  return TLM_OK;
}

int tlm_set_far_field_surfaces(tlm_handle* h, int32_t n, const tlm_far_field_surface* surfaces, const int32_t* rows,
                               const int32_t* sides, const int32_t* normals, const double* centres) {
  if (!h || n < 0 || (n && !surfaces)) return TLM_ERR_ARG;
  if (h->finalized) return fail(h, TLM_ERR_STATE, "tables are fixed after the first step");
  size_t faces = 0;
  for (int i = 0; i < n; i++) {
    const tlm_far_field_surface& f = surfaces[i];
    if (f.n_faces < 0 || f.n_theta < 1 || f.n_phi < 1 || (f.dim != 2 && f.dim != 3) ||
        (f.dim == 2 && f.n_theta != 1) || (double)f.n_theta * f.n_phi > 1e8)
      return fail(h, TLM_ERR_ARG, "far-field surface " + std::to_string(i + 1) + ": bad dimension or angles");
    faces += (size_t)f.n_faces;
  }
  if (faces && (!rows || !sides || !normals || !centres)) return TLM_ERR_ARG;
  for (size_t i = 0; i < faces; i++)
    if (sides[i] < 1 || sides[i] > 2 || normals[i] == 0 || normals[i] < -3 || normals[i] > 3)
      return fail(h, TLM_ERR_ARG, "far-field surface: bad face side or normal");
  h->ff.assign(surfaces, surfaces + n);
  h->ff_rows.assign(rows, rows + faces);
  h->ff_sides.assign(sides, sides + faces);
  h->ff_normals.assign(normals, normals + faces);
  h->ff_centres.assign(centres, centres + 3 * faces);
  return TLM_OK;
}

int tlm_read_far_field(tlm_handle* h, int32_t k, double* e) {
  if (!h || k < 0 || !e) return TLM_ERR_ARG;
  int rc;
  if ((rc = finalize(h))) return rc;
  if (k >= (int)h->ff.size()) return fail(h, TLM_ERR_ARG, "no such far-field surface");
  if ((rc = materialize(h))) return rc;  // a pending connect adds its currents
  const tlm_far_field_surface& f = h->ff[k];
  size_t item = (size_t)h->ff_item0;
  for (int i = 0; i < k; i++) item += 6 * (size_t)h->ff[i].n_faces;
  scn::FarFieldArgs a;
  // the reference forms j beta as j*cmplx(beta) (far_field_surfaces.F90:431,
  // :561): a default-kind CMPLX, so beta enters rounded to a float
  a.beta = (double)(float)(2.0 * scn::kPI * f.frequency / scn::kC0);
  a.dA = f.dim == 2 ? h->g.dl : h->g.dl * h->g.dl;
  a.theta_min = f.theta_min;
  a.theta_step = f.theta_step;
  a.phi_min = f.phi_min;
  a.phi_step = f.phi_step;
  a.n_theta = f.n_theta;
  a.n_phi = f.n_phi;
  a.dim = f.dim;
  a.nf = f.n_faces;
  // r0 = 1 m; exp(j beta r0) = (cos beta, sin beta)
  const double cb = cos(a.beta), sb = sin(a.beta);
  if (f.dim == 2) {
    // sqrt(j beta / (8 pi r0)) exp(j beta r0); the principal root of j y is
    // sqrt(y/2) (1 + j)
    const double rt = sqrt(0.5 * (a.beta / (8.0 * scn::kPI)));
    a.prop2_re = rt * cb - rt * sb;
    a.prop2_im = rt * sb + rt * cb;
  } else {
    // j beta exp(j beta r0) / (4 pi r0)
    a.prop2_re = (0.0 * cb - a.beta * sb) / (4.0 * scn::kPI);
    a.prop2_im = (0.0 * sb + a.beta * cb) / (4.0 * scn::kPI);
  }
  const int n_angles = f.n_theta * f.n_phi;
  hipLaunchKernelGGL(scn::k_far_field, dim3(n_angles), dim3(256), 0, h->stream, a, h->d_fo_acc + 2 * item,
                     h->d_ff_centres[k], h->d_ff_e);
  HIP_TRY(h, hipGetLastError());
  HIP_TRY(h, hipMemcpyAsync(e, h->d_ff_e, (size_t)n_angles * 4 * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  return TLM_OK;
}

int tlm_set_rcs_surface(tlm_handle* h, const tlm_rcs_surface* s, const int32_t* rows, const int32_t* sides,
                        const int32_t* normals, const double* centres) {
  if (!h) return TLM_ERR_ARG;
  if (h->finalized) return fail(h, TLM_ERR_STATE, "tables are fixed after the first step");
  if (!s) {
    h->rcs_on = false;
    return TLM_OK;
  }
  if (s->n_faces < 0 || s->n_points < 1 || s->points_per_cell < 1 || (s->n_faces && (!rows || !sides || !normals || !centres)))
    return fail(h, TLM_ERR_ARG, "RCS surface: bad faces or far-field points");
  for (int i = 0; i < s->n_faces; i++)
    if (sides[i] < 1 || sides[i] > 2 || normals[i] == 0 || normals[i] < -3 || normals[i] > 3)
      return fail(h, TLM_ERR_ARG, "RCS surface: bad face side or normal");
  h->rcs_on = true;
  h->rcs = *s;
  h->rcs_rows.assign(rows, rows + s->n_faces);
  h->rcs_sides.assign(sides, sides + s->n_faces);
  h->rcs_normals.assign(normals, normals + s->n_faces);
  h->rcs_centres.assign(centres, centres + 3 * (size_t)s->n_faces);
  return TLM_OK;
}

int tlm_read_rcs(tlm_handle* h, double* out) {
  if (!h || !out) return TLM_ERR_ARG;
  int rc;
  if ((rc = finalize(h))) return rc;
  if (!h->rcs_on) return fail(h, TLM_ERR_ARG, "no RCS surface");
  if ((rc = materialize(h))) return rc;
  HIP_TRY(h, hipMemcpyAsync(out, h->d_rcs, (size_t)h->rcs.n_points * 4 * sizeof(double), hipMemcpyDeviceToHost,
                            h->stream));
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  return TLM_OK;
}

int tlm_set_output_capacity(tlm_handle* h, int32_t steps) {
  if (!h) return TLM_ERR_ARG;
  if (steps < 1) return fail(h, TLM_ERR_ARG, "capacity must be positive");
  if (h->finalized) return fail(h, TLM_ERR_STATE, "capacity is fixed after the first step");
  h->cap = steps;
  return TLM_OK;
}

int tlm_upload_links(tlm_handle* h, const double* V) {
  if (!h || !V) return TLM_ERR_ARG;
  int rc;
  if ((rc = finalize(h))) return rc;
  size_t n = (size_t)12 * (h->nx + 2) * (h->ny + 2) * h->nz;
  void* tmp = nullptr;
  HIP_TRY(h, hipMalloc(&tmp, n * sizeof(double)));
  hipError_t e = hipMemcpyAsync(tmp, V, n * sizeof(double), hipMemcpyHostToDevice, h->stream);
  if (e == hipSuccess) {
    unsigned g = (unsigned)((h->P + 255) / 256);
    with_form(h->form, [&](auto t, auto) {
      using T = decltype(t);
      hipLaunchKernelGGL(scn::k_aos_to_planes<T>, dim3(g), dim3(256), 0, h->stream, h->d, (const double*)tmp,
                         (T*)h->planes[h->cur]);
    });
    e = hipStreamSynchronize(h->stream);
  }
  (void)hipFree(tmp);
  if (e != hipSuccess) return hip_fail(h, e, "upload_links");
  h->kind = Kind::INCIDENT;
  h->pend_connect = h->pend_boundary = false;
  return TLM_OK;
}

int tlm_download_links(tlm_handle* h, double* V) {
  if (!h || !V) return TLM_ERR_ARG;
  int rc;
  if ((rc = materialize(h))) return rc;
  size_t n = (size_t)12 * (h->nx + 2) * (h->ny + 2) * h->nz;
  void* tmp = nullptr;
  HIP_TRY(h, hipMalloc(&tmp, n * sizeof(double)));
  // keep the caller's halo: copy it in first, overwrite the interior
  hipError_t e = hipMemcpyAsync(tmp, V, n * sizeof(double), hipMemcpyHostToDevice, h->stream);
  if (e == hipSuccess) {
    unsigned g = (unsigned)((h->P + 255) / 256);
    with_form(h->form, [&](auto t, auto) {
      using T = decltype(t);
      hipLaunchKernelGGL(scn::k_planes_to_aos<T>, dim3(g), dim3(256), 0, h->stream, h->d,
                         (const T*)h->planes[h->cur], (double*)tmp);
    });
    e = hipMemcpyAsync(V, tmp, n * sizeof(double), hipMemcpyDeviceToHost, h->stream);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  (void)hipFree(tmp);
  if (e != hipSuccess) return hip_fail(h, e, "download_links");
  return TLM_OK;
}

int tlm_download_links_noalloc(tlm_handle* h, double* V) {
  if (!h || !V) return TLM_ERR_ARG;
  if (!h->finalized)
    return fail(h, TLM_ERR_STATE, "the allocation-free download needs a handle whose tables are set up "
                                  "(step, upload or tlm_host_cells first)");
  const size_t plane = (size_t)12 * (h->nx + 2) * (h->ny + 2);
  if (plane > kStagingDoubles)
    return fail(h, TLM_ERR_ARG, "one z-plane of the link array (" + std::to_string(plane) +
                                    " doubles) exceeds the allocation-free staging buffer (" +
                                    std::to_string(kStagingDoubles) + " doubles); use tlm_download_links");
  int rc;
  if ((rc = materialize(h))) return rc;  // a pending connect, on the existing buffers
  double* st = nullptr;
  if ((rc = staging_buffer(h, &st))) return rc;
  const int zc = (int)std::min<size_t>((size_t)h->nz, kStagingDoubles / plane);
  for (int z0 = 0; z0 < h->nz; z0 += zc) {
    const int z1 = std::min(h->nz, z0 + zc);
    const size_t n = plane * (size_t)(z1 - z0);
    // the caller's halo is kept: its slab goes in first, the kernel
    // overwrites the interior cells
    HIP_TRY(h, hipMemcpyAsync(st, V + plane * z0, n * sizeof(double), hipMemcpyHostToDevice, h->stream));
    const long long cells = (long long)h->nx * h->ny * (z1 - z0);
    const dim3 g((unsigned)((cells + 255) / 256));
    with_form(h->form, [&](auto t, auto) {
      using T = decltype(t);
      hipLaunchKernelGGL(scn::k_planes_to_aos_slab<T>, g, dim3(256), 0, h->stream, h->d,
                         (const T*)h->planes[h->cur], st, z0, z1);
    });
    HIP_TRY(h, hipGetLastError());
    HIP_TRY(h, hipMemcpyAsync(V + plane * z0, st, n * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  }
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  return TLM_OK;
}

int tlm_excitation(tlm_handle* h, int32_t timestep) {
  if (!h) return TLM_ERR_ARG;
  if (timestep < 1 || timestep > h->g.n_timesteps) return fail(h, TLM_ERR_ARG, "timestep out of range");
  // Excitation values are evaluated inside the scatter and connect kernels
  // from the function tables at their own timestep.
  return finalize(h);
}

int tlm_scatter(tlm_handle* h, int32_t timestep) {
  if (!h) return TLM_ERR_ARG;
  return do_scatter(h, timestep);
}

int tlm_connect(tlm_handle* h, int32_t timestep) {
  if (!h) return TLM_ERR_ARG;
  int rc;
  if ((rc = finalize(h))) return rc;
  if (h->kind != Kind::SCATTERED || h->pend_connect)
    return fail(h, TLM_ERR_STATE, "connect must follow a scatter");
  if (timestep < 1 || timestep > h->g.n_timesteps) return fail(h, TLM_ERR_ARG, "timestep out of range");
  if (h->n_host_faces && !h->hf_ready)
    return fail(h, TLM_ERR_STATE, "host-path faces need tlm_scatter_host_faces between scatter and connect");
  h->hf_ready = false;  // the pending connect owns them until the next pull
  h->pend_connect = true;
  h->pend_ts = timestep;
  return TLM_OK;
}

int tlm_outer_boundary(tlm_handle* h) {
  if (!h) return TLM_ERR_ARG;
  if (!h->pend_connect || h->pend_boundary)
    return fail(h, TLM_ERR_STATE, "outer_boundary must follow a connect");
  h->pend_boundary = true;
  return TLM_OK;
}

int tlm_host_cells(tlm_handle* h, int32_t* n_host, int32_t* cells) {
  if (!h || !n_host) return TLM_ERR_ARG;
  int rc;
  if ((rc = finalize(h))) return rc;
  *n_host = h->n_host;
  if (cells && h->n_host) std::copy(h->host_xyz.begin(), h->host_xyz.end(), cells);
  return TLM_OK;
}

// Host-path cells are handed over between a step's scatter and its connect:
// the buffer then holds their incident voltages (the scatter left them).
static int host_window(tlm_handle* h) {
  int rc;
  if ((rc = finalize(h))) return rc;
  if (h->kind != Kind::SCATTERED || h->pend_connect)
    return fail(h, TLM_ERR_STATE, "host-path cells are exchanged between tlm_scatter and tlm_connect");
  return TLM_OK;
}

int tlm_gather_host_cells(tlm_handle* h, double* V) {
  if (!h || !V) return TLM_ERR_ARG;
  int rc;
  if ((rc = host_window(h))) return rc;
  if (!h->n_host) return TLM_OK;
  const dim3 g((h->n_host + 255) / 256);
  with_form(h->form, [&](auto t, auto) {
    using T = decltype(t);
    hipLaunchKernelGGL(scn::k_gather_host<T>, g, dim3(256), 0, h->stream, (const T*)h->planes[h->cur], h->P,
                       h->d_host_cells, h->n_host, h->d_host_buf);
  });
  HIP_TRY(h, hipGetLastError());
  HIP_TRY(h, hipMemcpyAsync(V, h->d_host_buf, (size_t)12 * h->n_host * sizeof(double), hipMemcpyDeviceToHost,
                            h->stream));
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  return TLM_OK;
}

int tlm_scatter_host_cells(tlm_handle* h, const double* V) {
  if (!h || !V) return TLM_ERR_ARG;
  int rc;
  if ((rc = host_window(h))) return rc;
  if (!h->n_host) return TLM_OK;
  HIP_TRY(h, hipMemcpyAsync(h->d_host_buf, V, (size_t)12 * h->n_host * sizeof(double), hipMemcpyHostToDevice,
                            h->stream));
  const dim3 g((h->n_host + 255) / 256);
  with_form(h->form, [&](auto t, auto) {
    using T = decltype(t);
    hipLaunchKernelGGL(scn::k_scatter_host<T>, g, dim3(256), 0, h->stream, (T*)h->planes[h->cur], h->P,
                       h->d_host_cells, h->n_host, h->d_host_buf);
  });
  HIP_TRY(h, hipGetLastError());
  // the caller's array may be reused at once: wait for the copy
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  return TLM_OK;
}

int tlm_host_faces(tlm_handle* h, int32_t* n_host_faces, int32_t* faces) {
  if (!h || !n_host_faces) return TLM_ERR_ARG;
  int rc;
  if ((rc = finalize(h))) return rc;
  *n_host_faces = h->n_host_faces;
  if (faces && h->n_host_faces) std::copy(h->host_face_xyz.begin(), h->host_face_xyz.end(), faces);
  return TLM_OK;
}

int tlm_gather_host_faces(tlm_handle* h, double* V) {
  if (!h || !V) return TLM_ERR_ARG;
  int rc;
  if ((rc = host_window(h))) return rc;
  if (!h->n_host_faces) return TLM_OK;
  const dim3 g((h->n_host_faces + 255) / 256);
  with_form(h->form, [&](auto t, auto) {
    using T = decltype(t);
    hipLaunchKernelGGL(scn::k_gather_host_faces<T>, g, dim3(256), 0, h->stream, (const T*)h->planes[h->cur], h->P,
                       h->d_hf_cp, h->d_hf_cm, h->d_hf_ax, h->n_host_faces, h->d_hf_buf);
  });
  HIP_TRY(h, hipGetLastError());
  HIP_TRY(h, hipMemcpyAsync(V, h->d_hf_buf, (size_t)4 * h->n_host_faces * sizeof(double), hipMemcpyDeviceToHost,
                            h->stream));
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  return TLM_OK;
}

int tlm_scatter_host_faces(tlm_handle* h, const double* T) {
  if (!h || !T) return TLM_ERR_ARG;
  int rc;
  if ((rc = host_window(h))) return rc;
  if (h->n_host_faces)
    HIP_TRY(h, hipMemcpyAsync(h->d_hf_tot, T, (size_t)4 * h->n_host_faces * sizeof(double), hipMemcpyHostToDevice,
                              h->stream));
  // the caller's array may be reused at once: wait for the copy
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  h->hf_ready = true;
  return TLM_OK;
}

int tlm_step(tlm_handle* h, int32_t first_timestep, int32_t n_steps) {
  if (!h) return TLM_ERR_ARG;
  if (n_steps < 0 || first_timestep < 1 || first_timestep + n_steps - 1 > h->g.n_timesteps)
    return fail(h, TLM_ERR_ARG, "timestep range out of range");
  int rc;
  if ((rc = finalize(h))) return rc;
  if (h->n_host) return fail(h, TLM_ERR_STATE, "host-path cells need the split calls (tlm_scatter, tlm_connect)");
  if (h->n_host_faces)
    return fail(h, TLM_ERR_STATE, "host-path faces need the split calls (tlm_scatter, tlm_connect)");
  for (int t = first_timestep; t < first_timestep + n_steps; t++) {
    if ((rc = do_scatter(h, t))) return rc;
    h->pend_connect = true;
    h->pend_ts = t;
    h->pend_boundary = true;
  }
  return TLM_OK;
}

int tlm_synchronize(tlm_handle* h) {
  if (!h) return TLM_ERR_ARG;
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  return TLM_OK;
}

int tlm_read_cell_outputs(tlm_handle* h, double* out, int32_t max_steps, int32_t* n_read, int32_t* first_step) {
  if (!h || max_steps < 0 || (max_steps && !out)) return TLM_ERR_ARG;
  int rc;
  if ((rc = finalize(h))) return rc;
  if ((rc = drain_ring(h))) return rc;
  const int rows = h->d.n_out_rows;
  int n = std::min(h->host_steps, (int)max_steps);
  if (n > 0 && rows > 0) std::memcpy(out, h->host_out.data(), (size_t)n * rows * 6 * sizeof(double));
  if (n_read) *n_read = n;
  if (first_step) *first_step = h->host_first;
  if (n > 0) {
    h->host_out.erase(h->host_out.begin(), h->host_out.begin() + (size_t)n * rows * 6);
    h->host_steps -= n;
    h->host_first += n;
  }
  return TLM_OK;
}

int tlm_read_face_outputs(tlm_handle* h, double* out, int32_t max_steps, int32_t* n_read, int32_t* first_step) {
  if (!h || max_steps < 0 || (max_steps && !out)) return TLM_ERR_ARG;
  int rc;
  if ((rc = finalize(h))) return rc;
  if ((rc = materialize(h))) return rc;  // a pending connect computes its face outputs
  if ((rc = drain_face_ring(h))) return rc;
  const size_t w = (size_t)h->d.n_face_out_rows * 12;
  const int n = std::min(h->host_fsteps, (int)max_steps);
  if (n > 0 && w > 0) std::memcpy(out, h->host_fout.data(), (size_t)n * w * sizeof(double));
  if (n_read) *n_read = n;
  if (first_step) *first_step = h->host_ffirst;
  if (n > 0) {
    h->host_fout.erase(h->host_fout.begin(), h->host_fout.begin() + (size_t)n * w);
    h->host_fsteps -= n;
    h->host_ffirst += n;
  }
  return TLM_OK;
}

int tlm_read_dft(tlm_handle* h, double* re_im) {
  if (!h || !re_im) return TLM_ERR_ARG;
  int rc;
  if ((rc = finalize(h))) return rc;
  size_t n = h->pt_row.size() * h->freq.size() * 2;
  if (n) {
    HIP_TRY(h, hipMemcpyAsync(re_im, h->d_dft, n * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(h, hipStreamSynchronize(h->stream));
  }
  return TLM_OK;
}

int tlm_set_layout(tlm_handle* h, int32_t layout) {
  if (!h) return TLM_ERR_ARG;
  if (layout != TLM_LAYOUT_X_SEGMENTS && layout != TLM_LAYOUT_ROWS) return fail(h, TLM_ERR_ARG, "bad layout");
  if (h->finalized) return fail(h, TLM_ERR_STATE, "the layout is fixed once the device tables exist");
  if (layout != TLM_LAYOUT_ROWS && h->sched_variant >= 0 && h->sched_variant != 3)
    return fail(h, TLM_ERR_STATE, "the tile schedules need TLM_LAYOUT_ROWS");
  h->layout = layout;
  return TLM_OK;
}

int tlm_layout_segments(tlm_handle* h, int32_t* n_segments, int32_t* x0) {
  if (!h || !n_segments) return TLM_ERR_ARG;
  int rc;
  if ((rc = finalize(h))) return rc;
  *n_segments = h->d.nseg;
  if (x0)
    for (int s = 0; s <= h->d.nseg; s++) x0[s] = h->d.seg_x0[s];
  return TLM_OK;
}

int tlm_get_schedule(const tlm_handle* h, int32_t* variant, int32_t* rows, int32_t* zc, int32_t* xcd_remap,
                     int32_t* mixed, int32_t* streams) {
  if (!h) return TLM_ERR_ARG;
  if (variant) *variant = sched_variant(h);
  if (rows) *rows = h->sched_by;
  if (zc) *zc = h->sched_zc;
  if (xcd_remap) *xcd_remap = h->sched_xcd;
  if (mixed) *mixed = h->mixed_variant >= 0 ? h->mixed_variant : 1;
  if (streams) *streams = h->streams;
  return TLM_OK;
}

int tlm_set_streams(tlm_handle* h, int32_t streams) {
  if (!h) return TLM_ERR_ARG;
  if (streams != 1 && streams != 2) return fail(h, TLM_ERR_ARG, "streams must be 1 or 2");
  h->streams = streams;
  return TLM_OK;
}

int tlm_set_schedule(tlm_handle* h, int32_t variant, int32_t rows, int32_t zc, int32_t xcd_remap, int32_t mixed) {
  if (!h) return TLM_ERR_ARG;
  if (mixed < -1 || mixed > 1) return fail(h, TLM_ERR_ARG, "bad mixed-line kernel");
  if (variant < 0 || variant > 3 || (rows != 1 && rows != 2 && rows != 4 && (variant != 3 || (rows != 8 && rows != 16))) ||
      zc < 1)
    return fail(h, TLM_ERR_ARG, "bad bulk schedule");
  if (variant != 3 && h->layout != TLM_LAYOUT_ROWS)
    return fail(h, TLM_ERR_ARG, "the tile schedules (variants 0-2) need TLM_LAYOUT_ROWS");
  if (xcd_remap < 0 || xcd_remap > 2 || (xcd_remap == 2 && variant != 3)) return fail(h, TLM_ERR_ARG, "bad XCD mapping");
  if (ext_form(h) && (variant != 3 || rows != 1 || xcd_remap == 2 || mixed == 0))
    return fail(h, TLM_ERR_UNSUPPORTED, "the extended form runs the chunk kernel (variant 3, rows 1, XCD mapping 0 or "
                                        "1) and the LDS-staged mixed-line kernel only");
  h->mixed_variant = mixed;
  h->sched_variant = variant;
  h->sched_by = rows;
  h->sched_zc = zc;
  h->sched_xcd = xcd_remap;
  return TLM_OK;
}

int tlm_autotune(tlm_handle* h, int32_t first_timestep, int32_t steps_per_candidate, int32_t* steps_done) {
  if (!h || steps_per_candidate < 1 || first_timestep < 1) return TLM_ERR_ARG;
  int rc;
  if ((rc = finalize(h))) return rc;
  struct Cand { int v, rows, zc, xcd, mixed, streams; };
  std::vector<Cand> cand;
  for (int st = 2; st >= 1; st--)
    for (int xcd = 0; xcd < 3; xcd += 2)
      for (int mixed = 1; mixed >= 0; mixed--)
        if (!ext_form(h) || (xcd == 0 && mixed == 1)) cand.push_back({3, 1, 1, xcd, mixed, st});
  if (h->layout == TLM_LAYOUT_ROWS && !ext_form(h)) {
    const Cand tiles[] = {{0, 1, 1, 0, 1, 2}, {0, 2, 1, 0, 1, 2}, {0, 4, 1, 0, 1, 2}, {2, 1, 1, 0, 1, 2},
                          {2, 4, 1, 0, 1, 2}, {1, 4, 8, 0, 1, 2}};
    cand.insert(cand.end(), std::begin(tiles), std::end(tiles));
  }
  const int total = (int)cand.size() * (steps_per_candidate + 1);
  if (first_timestep + total - 1 > h->g.n_timesteps)
    return fail(h, TLM_ERR_ARG, "not enough timesteps left for the autotuner");
  hipEvent_t e0, e1;
  HIP_TRY(h, hipEventCreate(&e0));
  HIP_TRY(h, hipEventCreate(&e1));
  int ts = first_timestep, best = 0;
  float best_ms = 1e30f;
  for (size_t k = 0; k < cand.size(); k++) {
    h->sched_variant = cand[k].v;
    h->sched_by = cand[k].rows;
    h->sched_zc = cand[k].zc;
    h->sched_xcd = cand[k].xcd;
    h->mixed_variant = cand[k].mixed;
    h->streams = cand[k].streams;
    if ((rc = tlm_step(h, ts, 1))) break;  // first launch of this schedule, untimed
    ts += 1;
    if ((rc = hip_ok(h, hipEventRecord(e0, h->stream)))) break;
    if ((rc = tlm_step(h, ts, steps_per_candidate))) break;
    ts += steps_per_candidate;
    if ((rc = hip_ok(h, hipEventRecord(e1, h->stream)))) break;
    if ((rc = hip_ok(h, hipEventSynchronize(e1)))) break;
    float ms = 0.f;
    if ((rc = hip_ok(h, hipEventElapsedTime(&ms, e0, e1)))) break;
    if (ms < best_ms) {
      best_ms = ms;
      best = (int)k;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  h->sched_variant = cand[best].v;
  h->sched_by = cand[best].rows;
  h->sched_zc = cand[best].zc;
  h->sched_xcd = cand[best].xcd;
  h->mixed_variant = cand[best].mixed;
  h->streams = cand[best].streams;
  if (steps_done) *steps_done = ts - first_timestep;
  return TLM_OK;
}

int tlm_set_small_to_zero(tlm_handle* h, int32_t n_small_to_zero, double small_to_zero_value) {
  if (!h) return TLM_ERR_ARG;
  if (n_small_to_zero < 0 || !(small_to_zero_value >= 0.0)) return fail(h, TLM_ERR_ARG, "bad small_to_zero data");
  h->stz_n = n_small_to_zero;
  h->stz_value = small_to_zero_value;
  return TLM_OK;
}

int tlm_set_profiling(tlm_handle* h, int32_t enable) {
  if (!h) return TLM_ERR_ARG;
  int rc;
  if ((rc = resolve_profile(h))) return rc;
  h->profiling = enable != 0;
  return TLM_OK;
}

int tlm_read_kernel_times(tlm_handle* h, int64_t* launches, double* ms) {
  if (!h || !launches || !ms) return TLM_ERR_ARG;
  int rc;
  if ((rc = resolve_profile(h))) return rc;
  for (int k = 0; k < TLM_N_KCLASS; k++) {
    launches[k] = h->prof_launches[k];
    ms[k] = h->prof_ms[k];
    h->prof_launches[k] = 0;
    h->prof_ms[k] = 0.0;
  }
  return TLM_OK;
}

int tlm_kernel_bytes(tlm_handle* h, int64_t* cells, int64_t* bytes) {
  if (!h || !cells || !bytes) return TLM_ERR_ARG;
  int rc;
  if ((rc = finalize(h))) return rc;
  const int64_t sl = (int64_t)link_bytes(h);
  const int64_t L = h->d.line_len;
  const int64_t n_lines = h->d.seg_base[h->d.nseg] / L;
  // filter words of a dispersive cell: 3 components x (carried words read
  // and written) for eps and mu, FP64
  auto filter_bytes = [&](int m) -> int64_t { return filter_bytes_of(h, m); };
  // PML: links, the previous step's incident voltages pulled again (12
  // links), the nine streamed state words read and written (FP64 plain,
  // 48-bit in the reduced forms), packed cell word and 16-bit parameter row;
  // a cell with a special face reads and writes all its state words (24
  // FP64 plain; 9 FP64 and the nine streamed ones in the reduced forms)
  // instead of the second pull.
  const bool plain = h->form == TLM_FORM_PLAIN_FP64;
  const int64_t sw_bytes = plain ? 8 : 6;  // a streamed word: FP64, or 48 bits in the reduced forms
  const int64_t pml_rec_bytes = 36 * sl + 2 * sw_bytes * scn::SW_N + 4 + 2;
  const int64_t pml_fd_bytes =
      24 * sl + 2 * (plain ? 8 * scn::PS_N : 8 * scn::PC_N + sw_bytes * scn::SW_N) + 4 + 2;
  auto pml_bytes = [&](int64_t n, int64_t n_fd) { return (n - n_fd) * pml_rec_bytes + n_fd * pml_fd_bytes; };
  // generic: links, 7 int32 list fields, filter words
  std::vector<int32_t> gm(h->n_gen);
  if (h->n_gen) HIP_TRY(h, hipMemcpy(gm.data(), h->d.gen_mat, gm.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
  auto gen_bytes = [&](int i0, int i1) {
    int64_t b = 0;
    for (int i = i0; i < i1; i++) b += 24 * sl + 7 * 4 + filter_bytes(gm[i]);
    return b;
  };
  // material: links, cell word, 16-bit material number, filter number, filter words
  std::vector<uint16_t> mm(h->n_mat);
  if (h->n_mat) HIP_TRY(h, hipMemcpy(mm.data(), h->d.mat_num, mm.size() * sizeof(uint16_t), hipMemcpyDeviceToHost));
  auto mat_bytes = [&](int i0, int i1) {
    int64_t b = 0;
    for (int i = i0; i < i1; i++) b += 24 * sl + 4 + 2 + 4 + filter_bytes(mm[i]);
    return b;
  };
  const int64_t n_special_mixed = (h->n_pml - h->n_pure_pml) + (h->n_mat - h->n_pure_mat) + (h->n_gen - h->n_pure_gen);
  // bulk: cells of the pure bulk lines, 12 links read + 12 written, plus the
  // line table (line kernel) or the class mask (tile kernels)
  const int64_t n_bulk =
      h->N - h->n_mixed_cells - h->n_pure_pml - h->n_pure_mat - h->n_pure_gen - h->n_pm - h->n_matline_cells;
  cells[TLM_KCLASS_BULK] = n_bulk;
  // direct line kernel: one 16-byte chunk record per 64 cells
  const bool direct = sched_variant(h) == 3 && h->sched_by == 1 && h->sched_xcd != 2;
  bytes[TLM_KCLASS_BULK] = n_bulk * 24 * sl +
                           (direct ? 16 * (int64_t)h->n_bulk_chunks
                                   : (sched_variant(h) == 3 ? 4 * n_lines : (h->N + 7) / 8));
  // mixed lines: every cell with its class's bytes, plus two slot words
  cells[TLM_KCLASS_MIXED] = h->n_mixed_cells;
  bytes[TLM_KCLASS_MIXED] = (h->n_mixed_cells - n_special_mixed) * 24 * sl +
                            pml_bytes(h->n_pml - h->n_pure_pml, h->n_pml_fd_all - h->n_pml_fd_pure) +
                            mat_bytes(h->n_pure_mat, h->n_mat) + gen_bytes(h->n_pure_gen, h->n_gen) +
                            (int64_t)h->n_mixed_blocks * 256 * 8;
  // pure PML lines run from chunk records: no cell word, a 32-byte record
  // per chunk
  cells[TLM_KCLASS_PML] = h->n_pure_pml;
  bytes[TLM_KCLASS_PML] = pml_bytes(h->n_pure_pml, h->n_pml_fd_pure) - 4 * (int64_t)h->n_pure_pml +
                          32 * (int64_t)h->n_pml_chunks;
  cells[TLM_KCLASS_GENERIC] = h->n_pure_gen;
  bytes[TLM_KCLASS_GENERIC] = gen_bytes(0, h->n_pure_gen);
  // material class: pure material lines from the list, material lines from
  // records (links of every cell, filter words of the material cells, 32
  // bytes per record)
  cells[TLM_KCLASS_MATERIAL] = h->n_pure_mat + h->n_matline_cells;
  bytes[TLM_KCLASS_MATERIAL] = mat_bytes(0, h->n_pure_mat) + h->n_matline_cells * 24 * sl +
                               h->matline_filter_bytes + 32 * (int64_t)h->n_mat_recs;
  // PML cells holding a material: links, PM_N FP64 state words read and
  // written, five int32 list fields
  cells[TLM_KCLASS_PML_MATERIAL] = h->n_pm;
  bytes[TLM_KCLASS_PML_MATERIAL] = (int64_t)h->n_pm * (24 * sl + 2 * 8 * scn::PM_N + 5 * 4);
  cells[TLM_KCLASS_DFT] = (int64_t)h->pt_row.size() * (int64_t)h->freq.size();
  bytes[TLM_KCLASS_DFT] = cells[TLM_KCLASS_DFT] * 2 * 2 * 8;
  return TLM_OK;
}

}  // extern "C"
