// Host side of the batched solver: packs a batch of packing LPs into the
// device layout, splits it into size classes (basis inverse in LDS or in a
// global scratch slice), launches one persistent kernel per class (captured in
// a hipGraph), and tunes the launch shape.
#include "ds_batch_host.hpp"

#include "ds_warm.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>

namespace pds {

namespace {
// Dynamic LDS per workgroup: 160 KB per CU, less the static reduction buffers.
constexpr size_t LDS_BUDGET = 160 * 1024 - 2048;
double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
template <typename T>
void up(DBuf<T>& b, const std::vector<T>& v, hipStream_t s) {
  b.alloc(std::max<size_t>(v.size(), 1));
  if (!v.empty()) PDS_HIP(hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
}
}  // namespace

BatchSolver::BatchSolver(int device) : device_(device) {
  PDS_HIP(hipSetDevice(device_));
  PDS_HIP(hipStreamCreateWithFlags(&st_, hipStreamNonBlocking));
  PDS_HIP(hipDeviceGetAttribute(&cus_, hipDeviceAttributeMultiprocessorCount, device_));
}

BatchSolver::~BatchSolver() {
  drop_graph();
  if (st_) {
    (void)hipStreamSynchronize(st_);
    (void)hipStreamDestroy(st_);
  }
}

void BatchSolver::drop_graph() {
  if (exec_) (void)hipGraphExecDestroy(exec_);
  if (graph_) (void)hipGraphDestroy(graph_);
  exec_ = nullptr;
  graph_ = nullptr;
}

void BatchSolver::load(int count, const int* m, const int* n, const long long* nnz_off, const int* col_ptr,
                       const int* row_idx, const double* val, const double* b, const double* c) {
  if (count <= 0) throw std::invalid_argument("empty batch");
  count_ = count;
  hm_.assign(m, m + count);
  hn_.assign(n, n + count);
  std::vector<long long> cp_off(count), rp_off(count), nz(count), mo(count), vo(count);
  long long cpo = 0, rpo = 0, mso = 0, vso = 0, nso = 0;
  mmax_all_ = nmax_all_ = 0;
  for (int k = 0; k < count; ++k) {
    if (m[k] <= 0 || n[k] <= 0) throw std::invalid_argument("LP with no rows or columns");
    cp_off[k] = cpo;
    rp_off[k] = rpo;
    nz[k] = nnz_off[k];
    mo[k] = mso;
    vo[k] = vso;
    cpo += n[k] + 1;
    rpo += m[k] + 1;
    mso += m[k];
    vso += n[k] + m[k];
    nso += n[k];
    mmax_all_ = std::max(mmax_all_, m[k]);
    nmax_all_ = std::max(nmax_all_, n[k]);
  }
  const long long nnz = nnz_off[count];
  sum_m_ = mso;
  sum_v_ = vso;
  std::vector<int> cpv(col_ptr, col_ptr + cpo), rowp(rpo), coli(nnz);
  std::vector<int> riv(row_idx, row_idx + nnz);
  std::vector<double> cvv(val, val + nnz), rvv(nnz);
  std::vector<double> cc(vso), lv(vso), uv(vso);
  std::vector<int8_t> btv(vso);
  long long bo = 0, co = 0;
  for (int k = 0; k < count; ++k) {
    const int mk = m[k], nk = n[k];
    const int* cpk = col_ptr + cp_off[k];
    const long long z0 = nnz_off[k];
    if (cpk[0] != 0 || cpk[nk] != (int)(nnz_off[k + 1] - z0)) throw std::invalid_argument("col_ptr / nnz_off mismatch");
    // CSR of LP k, entries of a row in ascending column order.
    int* rp = rowp.data() + rp_off[k];
    std::fill(rp, rp + mk + 1, 0);
    for (int e = 0; e < cpk[nk]; ++e) {
      int r = row_idx[z0 + e];
      if (r < 0 || r >= mk) throw std::invalid_argument("row index out of range");
      rp[r + 1]++;
    }
    for (int r = 0; r < mk; ++r) rp[r + 1] += rp[r];
    std::vector<int> fill(rp, rp + mk);
    for (int j = 0; j < nk; ++j)
      for (int e = cpk[j]; e < cpk[j + 1]; ++e) {
        int r = row_idx[z0 + e];
        coli[z0 + fill[r]] = j;
        rvv[z0 + fill[r]] = val[z0 + e];
        fill[r]++;
      }
    // Computational form: structurals x >= 0 (Lower), slack of row i in
    // [-b_i, +inf) (Lower), costs (c, 0).
    for (int j = 0; j < nk; ++j) {
      cc[vo[k] + j] = c[co + j];
      lv[vo[k] + j] = 0.0;
      uv[vo[k] + j] = std::numeric_limits<double>::infinity();
      btv[vo[k] + j] = BT_LOWER;
    }
    for (int i = 0; i < mk; ++i) {
      cc[vo[k] + nk + i] = 0.0;
      lv[vo[k] + nk + i] = -b[bo + i];
      uv[vo[k] + nk + i] = std::numeric_limits<double>::infinity();
      btv[vo[k] + nk + i] = BT_LOWER;
    }
    bo += mk;
    co += nk;
  }
  up(m_, hm_, st_);
  up(n_, hn_, st_);
  up(cp_off_, cp_off, st_);
  up(rp_off_, rp_off, st_);
  std::vector<long long> nzo(nnz_off, nnz_off + count);
  up(nz_off_, nzo, st_);
  up(mo_, mo, st_);
  up(vo_, vo, st_);
  up(colptr_, cpv, st_);
  up(rowidx_, riv, st_);
  up(cval_, cvv, st_);
  up(rowptr_, rowp, st_);
  up(colidx_, coli, st_);
  up(rval_, rvv, st_);
  up(c_, cc, st_);
  up(l_, lv, st_);
  up(u_, uv, st_);
  up(bt_, btv, st_);
  status_.alloc(count);
  obj_.alloc(count);
  iters_.alloc(count);
  basis_.alloc(mso);
  x_.alloc(vso);
  next_.alloc(4);
  repairs_.alloc(count);
  solved_ = false;
  PDS_HIP(hipMemsetAsync(repairs_.p, 0, sizeof(int) * count, st_));
  warm_on_ = false;
  PDS_HIP(hipStreamSynchronize(st_));
  drop_graph();
}

void BatchSolver::set_warm(const int* basis) {
  if (!basis) {
    warm_on_ = false;
    drop_graph();
    return;
  }
  std::vector<int> w(basis, basis + sum_m_);
  // Every starting basis must list m_k distinct variables of its LP.
  long long o = 0;
  for (int k = 0; k < count_; ++k) {
    const int N = hm_[k] + hn_[k];
    std::vector<char> seen(N, 0);
    for (int i = 0; i < hm_[k]; ++i) {
      int j = w[o + i];
      if (j < 0 || j >= N || seen[j]) throw std::invalid_argument("invalid starting basis for LP " + std::to_string(k));
      seen[j] = 1;
    }
    o += hm_[k];
  }
  up(warm_, w, st_);
  PDS_HIP(hipStreamSynchronize(st_));
  warm_on_ = true;
  drop_graph();
}

// Bases found elsewhere (for instance by the single-LP solver, for LPs too large
// for the batched kernel) stand in for a solve of the batch: they become the
// labels of the learned warm start.
void BatchSolver::set_bases(const int* basis) {
  std::vector<int> w(basis, basis + sum_m_);
  long long o = 0;
  for (int k = 0; k < count_; ++k) {
    const int N = hm_[k] + hn_[k];
    std::vector<char> seen(N, 0);
    for (int i = 0; i < hm_[k]; ++i) {
      int j = w[o + i];
      if (j < 0 || j >= N || seen[j]) throw std::invalid_argument("invalid basis for LP " + std::to_string(k));
      seen[j] = 1;
    }
    o += hm_[k];
  }
  up(basis_, w, st_);
  PDS_HIP(hipStreamSynchronize(st_));
  solved_ = true;
}

// Build the per-class launch descriptions for the LP subset `items`.
void BatchSolver::plan(const std::vector<int>& items, const ds_batch_params& p) {
  std::vector<int> small, large;
  for (int k : items) (hm_[k] <= p.lds_inv_max_m ? small : large).push_back(k);
  auto by_size = [&](int a, int b) {
    long long wa = (long long)hm_[a] * hm_[a], wb = (long long)hm_[b] * hm_[b];
    return wa != wb ? wa > wb : a < b;
  };
  std::sort(small.begin(), small.end(), by_size);
  std::sort(large.begin(), large.end(), by_size);
  std::vector<int> order;
  order.insert(order.end(), large.begin(), large.end());
  order.insert(order.end(), small.begin(), small.end());
  up(order_, order, st_);
  classes_.clear();
  int at = 0;
  for (int cls = 0; cls < 2; ++cls) {
    const std::vector<int>& v = cls == 0 ? large : small;
    if (v.empty()) continue;
    Class C;
    C.first = at;
    C.count = (int)v.size();
    C.inv_in_lds = cls == 1;
    C.mmax = C.nmax = 0;
    for (int k : v) {
      C.mmax = std::max(C.mmax, hm_[k]);
      C.nmax = std::max(C.nmax, hn_[k]);
    }
    C.grid = std::min(C.count, std::max(1, cus_ * std::max(1, p.wgs_per_cu)));
    C.counter = cls;
    classes_.push_back(C);
    at += C.count;
  }
  size_t scratch = 0;
  for (const Class& C : classes_)
    if (!C.inv_in_lds) scratch = std::max(scratch, (size_t)C.grid * C.mmax * C.mmax);
  if (scratch) gscratch_.alloc(scratch);
  PDS_HIP(hipStreamSynchronize(st_));
}

BatchDev BatchSolver::make_dev() const {
  BatchDev D;
  D.count = count_;
  D.m = m_.p;
  D.n = n_.p;
  D.cp_off = cp_off_.p;
  D.rp_off = rp_off_.p;
  D.nz_off = nz_off_.p;
  D.mo = mo_.p;
  D.vo = vo_.p;
  D.colptr = colptr_.p;
  D.rowidx = rowidx_.p;
  D.cval = cval_.p;
  D.rowptr = rowptr_.p;
  D.colidx = colidx_.p;
  D.rval = rval_.p;
  D.c = c_.p;
  D.l = l_.p;
  D.u = u_.p;
  D.bt = bt_.p;
  D.order = nullptr;
  D.order_count = 0;
  D.warm = warm_on_ ? warm_.p : nullptr;
  D.repair_tol = warm_on_ ? repair_tol_ : 0.0;
  D.repairs = repairs_.p;
  D.status = status_.p;
  D.obj = obj_.p;
  D.iters = iters_.p;
  D.basis_out = basis_.p;
  D.x_out = x_.p;
  D.next = nullptr;
  D.gscratch = gscratch_.p;
  D.mmax = mmax_all_;
  D.nmax = nmax_all_;
  D.inv_in_lds = 0;
  D.gj_nb = 1;
  D.refactor_freq = 150;
  D.max_iter = 0;
  D.perturb = 0;
  D.cleanup = 0;
  D.seed = 0;
  return D;
}

void BatchSolver::enqueue(const ds_batch_params& p) {
  for (const Class& C : classes_) {
    BatchDev D = make_dev();
    D.order = order_.p + C.first;
    D.order_count = C.count;
    D.next = next_.p + C.counter;
    D.gscratch = gscratch_.p;
    D.mmax = C.mmax;
    D.nmax = C.nmax;
    D.inv_in_lds = C.inv_in_lds ? 1 : 0;
    D.gj_nb = C.gj_nb;
    D.refactor_freq = p.refactor_freq > 0 ? p.refactor_freq : 150;
    D.max_iter = p.max_iterations;
    D.perturb = p.perturb;
    D.cleanup = p.cleanup;
    D.seed = p.perturb_seed;
    launch_batch_reset(st_, D.next);
    launch_batch(st_, D, C.grid, p.block);
  }
}

void BatchSolver::features(float* F_dev) {
  BatchDev D = make_dev();
  launch_features(st_, D, count_, mmax_all_, nmax_all_, F_dev);
  PDS_HIP(hipGetLastError());
}

void BatchSolver::labels(float* y_dev) {
  if (!solved_) throw std::logic_error("labels need a solved batch");
  BatchDev D = make_dev();
  launch_labels(st_, D, count_, basis_.p, y_dev);
  PDS_HIP(hipGetLastError());
}

void BatchSolver::select_warm(const float* z_dev) {
  BatchDev D = make_dev();
  warm_.alloc(std::max<long long>(sum_m_, 1));
  launch_select(st_, D, count_, mmax_all_ + nmax_all_, z_dev, warm_.p);
  PDS_HIP(hipGetLastError());
  warm_on_ = true;
  drop_graph();
}

double BatchSolver::run(const std::vector<int>& items, const ds_batch_params& p, bool use_graph) {
  plan(items, p);
  for (Class& C : classes_) {
    C.gj_nb = batch_panel_width(C.mmax, C.nmax, C.inv_in_lds, LDS_BUDGET);
    if (C.gj_nb == 0) throw std::invalid_argument("batch size class exceeds the LDS budget");
  }
  if (use_graph) {
    drop_graph();
    PDS_HIP(hipStreamBeginCapture(st_, hipStreamCaptureModeThreadLocal));
    enqueue(p);
    PDS_HIP(hipStreamEndCapture(st_, &graph_));
    PDS_HIP(hipGraphInstantiate(&exec_, graph_, nullptr, nullptr, 0));
  }
  PDS_HIP(hipStreamSynchronize(st_));
  hipEvent_t e0, e1;
  PDS_HIP(hipEventCreate(&e0));
  PDS_HIP(hipEventCreate(&e1));
  PDS_HIP(hipEventRecord(e0, st_));
  if (use_graph) PDS_HIP(hipGraphLaunch(exec_, st_));
  else enqueue(p);
  PDS_HIP(hipEventRecord(e1, st_));
  PDS_HIP(hipEventSynchronize(e1));
  PDS_HIP(hipGetLastError());
  float ms = 0.f;
  PDS_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return ms;
}

double BatchSolver::solve(const ds_batch_params& p) {
  std::vector<int> all(count_);
  for (int k = 0; k < count_; ++k) all[k] = k;
  const double ms = run(all, p, p.use_graph != 0);
  solved_ = true;
  return ms;
}

double BatchSolver::replay() {
  if (!exec_) throw std::logic_error("no captured batch graph");
  hipEvent_t e0, e1;
  PDS_HIP(hipEventCreate(&e0));
  PDS_HIP(hipEventCreate(&e1));
  PDS_HIP(hipEventRecord(e0, st_));
  PDS_HIP(hipGraphLaunch(exec_, st_));
  PDS_HIP(hipEventRecord(e1, st_));
  PDS_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  PDS_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return ms;
}

void BatchSolver::get_warm(int* out) {
  if (!warm_on_) throw std::logic_error("no starting bases set");
  PDS_HIP(hipMemcpyAsync(out, warm_.p, sizeof(int) * sum_m_, hipMemcpyDeviceToHost, st_));
  PDS_HIP(hipStreamSynchronize(st_));
}

void BatchSolver::repairs(int* out) {
  PDS_HIP(hipMemcpyAsync(out, repairs_.p, sizeof(int) * count_, hipMemcpyDeviceToHost, st_));
  PDS_HIP(hipStreamSynchronize(st_));
}

void BatchSolver::results(int* status, double* obj, int* iters, int* basis, double* x) {
  if (status) PDS_HIP(hipMemcpyAsync(status, status_.p, sizeof(int) * count_, hipMemcpyDeviceToHost, st_));
  if (obj) PDS_HIP(hipMemcpyAsync(obj, obj_.p, sizeof(double) * count_, hipMemcpyDeviceToHost, st_));
  if (iters) PDS_HIP(hipMemcpyAsync(iters, iters_.p, sizeof(int) * count_, hipMemcpyDeviceToHost, st_));
  if (basis) PDS_HIP(hipMemcpyAsync(basis, basis_.p, sizeof(int) * sum_m_, hipMemcpyDeviceToHost, st_));
  if (x) PDS_HIP(hipMemcpyAsync(x, x_.p, sizeof(double) * sum_v_, hipMemcpyDeviceToHost, st_));
  PDS_HIP(hipStreamSynchronize(st_));
}

// Times every candidate launch shape on an evenly spread sample of the batch
// and keeps the fastest; candidates must agree on status and objective.
void BatchSolver::tune(int sample, ds_batch_params& best, double& best_ms, int& tried) {
  std::vector<int> items;
  const int step = std::max(1, count_ / std::max(1, sample));
  for (int k = 0; k < count_ && (int)items.size() < sample; k += step) items.push_back(k);
  const int blocks[3] = {128, 256, 512};
  const int ldsm[3] = {0, 96, 128};
  const int occ[3] = {1, 2, 4};
  const int refac[3] = {150, 300, 600};
  std::vector<int> ref_st(count_);
  std::vector<double> ref_obj(count_);
  bool have_ref = false;
  best_ms = 1e300;
  tried = 0;
  // Every candidate must end each sampled LP with the same status and an
  // objective within 1e-9 of the first candidate's.
  auto trial = [&](const ds_batch_params& p) {
    double ms;
    try {
      ms = run(items, p, false);
    } catch (const std::invalid_argument&) {
      return;
    }
    ++tried;
    std::vector<int> st(count_);
    std::vector<double> ob(count_);
    results(st.data(), ob.data(), nullptr, nullptr, nullptr);
    if (!have_ref) {
      ref_st = st;
      ref_obj = ob;
      have_ref = true;
    } else {
      for (int k : items)
        if (st[k] != ref_st[k] || std::fabs(ob[k] - ref_obj[k]) > 1e-9 * std::max(1.0, std::fabs(ref_obj[k])))
          throw std::runtime_error("tuned launch shapes disagree on a result");
    }
    if (ms < best_ms) {
      best_ms = ms;
      best = p;
    }
  };
  // Stage 1: launch shape (workgroup size x LDS-inverse threshold x
  // occupancy) at the given refactor interval; stage 2: refactor interval.
  const ds_batch_params base = best;
  for (int bi = 0; bi < 3; ++bi)
    for (int li = 0; li < 3; ++li)
      for (int oi = 0; oi < 3; ++oi) {
        ds_batch_params p = base;
        p.block = blocks[bi];
        p.lds_inv_max_m = ldsm[li];
        p.wgs_per_cu = occ[oi];
        trial(p);
      }
  const ds_batch_params shape = best;
  for (int ri = 0; ri < 3; ++ri) {
    if (refac[ri] == shape.refactor_freq) continue;
    ds_batch_params p = shape;
    p.refactor_freq = refac[ri];
    trial(p);
  }
}

}  // namespace pds
