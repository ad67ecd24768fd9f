// Singles propagation entry points: the batched absorption plan (frame-major
// conveyor, plain per-sample baseline) and the drop-in symbols
// propagate_vec_coupling_S / propagate_vector / propagate_matrix.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <tuple>

#include "nise_internal.h"

using namespace nise;

namespace nise {
int dia_propagate_host(const t_non *non, const float *H, float *cr, float *ci, int sign, int nvec,
                       hipStream_t st);
int dia_s_propagate_host(const t_non *non, const float *H, float *cr, float *ci, int sign, int nvec,
                         hipStream_t st, int *elements);
}

// ---------------------------------------------------------------------------
// Absorption plan
// ---------------------------------------------------------------------------
struct nise_abs_plan {
  t_non non{};
  int N = 0, tmax = 0, m = 1, begin = 0, end = 0, sample = 1;
  int kernel = 0;
  int64_t f_lo = 0, f_hi = 0;  // frames used
  FrameSource src;
  const float *mu_dev = nullptr;
  int64_t mu_stride = 0;
  DevBuf H_own, mu_own, C_own, shift_site, shift_ab, pos_own, box_own;
  Schedule sch;                // active schedule (one of the two below)
  Schedule sch_asap, sch_rows; // layered (slot conveyor) / row schedule
  int rows = 0;                // rows of the row conveyor, 0: slot conveyor
  DevBuf ab, layer_off, diag, cs, slots, wg_begin, partial, out, steps, row_steps, row_sys, row_wmap;
  size_t row_smem = 0;
  bool row_pf = false;  // the row conveyor prefetches each frame's first site batch
  int nsteps = 0;
  DevBuf ws_cs, resp_plain;
  int vp = 16, spw = 5, nwg = 0, block = 256;
  int npart = 0;  // rows of partial responses (workgroups, or vectors for the row conveyor)
  // gradient runs: parameters and the adjoint's buffers (every == 0: off)
  DevBuf delta, jf, ck, gtab, part_d, part_c, grad, adj_slots, adj_wg, adj_steps;
  int every = 0, nck = 0, adj_vp = 0, adj_nwg = 0, adj_nsteps = 0;
  std::vector<uint8_t> mask;   // union coupling structure
  hipStream_t st = nullptr;
  hipEvent_t ev[4] = {};
  float last_ms[4] = {0, 0, 0, 0};
  bool graph_ready = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  ~nise_abs_plan() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    for (auto &e : ev)
      if (e) (void)hipEventDestroy(e);
    if (st) (void)hipStreamDestroy(st);
  }
};

namespace {

// Allocates at least 16 bytes so that empty inputs (no couplings) still get
// a valid device pointer; copies only when there is data.
int upload(DevBuf &dst, const void *src, size_t bytes, hipStream_t st) {
  // work queued on the plan's stream (an execute without a result fetch
  // returns early) may still read dst: the copy below does not wait for it
  if (st) NISE_HIP_CHECK(hipStreamSynchronize(st));
  int rc = dst.alloc(bytes < 16 ? 16 : bytes);
  if (rc) return rc;
  // synchronous: the source may be pageable memory owned by the caller
  if (bytes && src) NISE_HIP_CHECK(hipMemcpy(dst.p, src, bytes, hipMemcpyHostToDevice));
  return NISE_OK;
}

bool is_coupling_type(const t_non *non) { return std::strcmp(non->hamiltonian, "Coupling") == 0; }

// Descriptor of the plan's single system for the row conveyor (device
// copy and workgroup map; refreshed when checkpoints are switched on).
int upload_rows_desc(nise_abs_plan *P) {
  RowsArgs r;
  r.N = P->N, r.tmax = P->tmax, r.m = P->m, r.nlayers = P->sch.nlayers(), r.K = P->sch.K, r.rows = P->rows;
  r.f_lo = P->f_lo, r.nvec = 3 * (P->end - P->begin), r.begin = P->begin, r.sample = P->sample;
  r.steps = P->row_steps.as<uint2>(), r.cs = P->cs.as<float2>(), r.diag = P->diag.as<float2>();
  r.mu_rec = P->mu_dev, r.mu_stride = P->mu_stride, r.partial = P->partial.as<double2>();
  if (P->every) r.hist = P->ck.as<float2>(), r.hist_every = P->every, r.nck = P->nck;
  std::vector<int2> wmap;
  int rc = rows_prepare(&r, 1, P->rows, wmap, P->row_smem, P->row_pf);
  if (rc) return rc;
  if ((rc = upload(P->row_sys, &r, sizeof(r), P->st))) return rc;
  return upload(P->row_wmap, wmap.data(), wmap.size() * sizeof(int2), P->st);
}

// Row conveyor set-up: schedule, step table, tables and partials.
int build_rows(nise_abs_plan *P) {
  int rc;
  P->sch = P->sch_rows;
  std::vector<uint2> tab;
  rows_step_table(P->sch, P->rows, tab);
  if ((rc = upload(P->row_steps, tab.data(), tab.size() * sizeof(uint2), P->st))) return rc;
  std::vector<uint32_t> ab;  // the frame tables follow the step table's order
  rows_device_order(P->sch, P->rows, ab);
  if ((rc = upload(P->ab, ab.data(), ab.size() * 4, P->st))) return rc;
  if ((rc = upload(P->layer_off, P->sch.layer_off.data(), P->sch.layer_off.size() * 4, P->st))) return rc;
  const int64_t nfr = P->f_hi - P->f_lo;
  if ((rc = P->cs.alloc((size_t)nfr * std::max(P->sch.K, 1) * sizeof(float2)))) return rc;
  if (P->every && (rc = P->jf.alloc((size_t)nfr * std::max(P->sch.K, 1) * 4))) return rc;
  const int G = 64 / P->rows;
  P->nwg = (3 * (P->end - P->begin) + G - 1) / G;
  P->npart = 3 * (P->end - P->begin);  // one row per vector
  if ((rc = P->partial.alloc((size_t)std::max(P->npart, 1) * P->tmax * sizeof(double2)))) return rc;
  return upload_rows_desc(P);
}

// Slot assignment: consecutive samples, three polarisations each, spw
// samples per workgroup (3*spw <= vp slots).
int build_slots(nise_abs_plan *P) {
  if (P->rows) return build_rows(P);
  int rc0;
  P->sch = P->sch_asap;
  if ((rc0 = upload(P->ab, P->sch.ab.data(), P->sch.ab.size() * 4, P->st))) return rc0;
  if ((rc0 = upload(P->layer_off, P->sch.layer_off.data(), P->sch.layer_off.size() * 4, P->st))) return rc0;
  std::vector<SlotDesc> sl;
  std::vector<int> wb;
  int ns = P->end - P->begin;
  for (int s = 0; s < ns; ++s) {
    if (s % P->spw == 0) wb.push_back((int)sl.size());
    int fs = (P->begin + s) * P->sample;
    for (int x = 0; x < 3; ++x) {
      const int hist = P->every ? (int)(3 * (int64_t)s + x) * P->nck : -1;  // checkpoint rows
      sl.push_back(SlotDesc{fs, P->tmax, P->tmax - 1, x, 0, -1, x == 0 ? 3 : 0, hist});
    }
  }
  wb.push_back((int)sl.size());
  P->nwg = (int)wb.size() - 1;
  P->npart = P->nwg;
  int rc;
  std::vector<int2> stp;
  build_steps(P->sch, P->vp, P->block, stp);
  P->nsteps = (int)stp.size();
  if ((rc = upload(P->steps, stp.data(), stp.size() * sizeof(int2), P->st))) return rc;
  if ((rc = upload(P->slots, sl.data(), sl.size() * sizeof(SlotDesc), P->st))) return rc;
  if ((rc = upload(P->wg_begin, wb.data(), wb.size() * sizeof(int), P->st))) return rc;
  if ((rc = P->partial.alloc((size_t)std::max(P->nwg, 1) * P->tmax * sizeof(double2)))) return rc;
  return NISE_OK;
}

int run_tables(nise_abs_plan *P) {
  float fac = trotter_factor(P->non.deltat, 1, P->m);
  return launch_frame_tables(P->src, P->N, P->sch, P->ab.as<uint32_t>(), fac, P->f_lo,
                             P->f_hi - P->f_lo, P->diag.as<float2>(), P->cs.as<float2>(), P->st,
                             P->every ? P->jf.as<float>() : nullptr);
}

int run_conveyor(nise_abs_plan *P) {
  if (P->rows)
    return launch_rows(P->row_sys.as<RowsArgs>(), P->row_wmap.as<int2>(), P->nwg, P->rows, P->row_pf, P->row_smem,
                       P->st);
  ConveyorArgs a;
  a.N = P->N, a.K = P->sch.K, a.nlayers = P->sch.nlayers(), a.m = P->m;
  a.f_lo = P->f_lo;
  a.ab = P->ab.as<uint32_t>(), a.layer_off = P->layer_off.as<int>();
  a.diag = P->diag.as<float2>(), a.cs = P->cs.as<float2>();
  a.mu_rec = P->mu_dev, a.mu_stride = P->mu_stride;
  a.slots = P->slots.as<SlotDesc>(), a.wg_slot_begin = P->wg_begin.as<int>(), a.nwg = P->nwg;
  a.tmax_acc = P->tmax, a.partial = P->partial.as<double2>();
  a.steps = P->steps.as<int2>(), a.nsteps = P->nsteps;
  if (P->every) a.hist = P->ck.as<float2>(), a.hist_every = P->every;
  return launch_conveyor(a, P->vp, P->block, P->st);
}

int run_reduce(nise_abs_plan *P) {
  float *re = P->out.as<float>();
  return launch_reduce_partials(P->partial.as<double2>(), P->npart, P->tmax, re, re + P->tmax, P->st);
}

__global__ void zero_d2(double2 *p, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = make_double2(0, 0);
}
__global__ void d2_to_f(const double2 *p, int n, float *re, float *im) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) re[i] = (float)p[i].x, im[i] = (float)p[i].y;
}

int run_plain(nise_abs_plan *P) {
  double2 *resp = P->resp_plain.as<double2>();
  hipLaunchKernelGGL(zero_d2, dim3((P->tmax + 255) / 256), dim3(256), 0, P->st, resp, P->tmax);
  PlainArgs a;
  a.N = P->N, a.m = P->m, a.tmax = P->tmax;
  a.fac = trotter_factor(P->non.deltat, 1, P->m);
  a.src = P->src, a.mu_rec = P->mu_dev, a.mu_stride = P->mu_stride;
  a.resp = resp, a.ws_cs = P->ws_cs.as<float2>();
  a.ab = P->ab.as<uint32_t>(), a.layer_off = P->layer_off.as<int>();
  a.K = P->sch_asap.K, a.nlayers = P->sch_asap.nlayers();
  a.cs_lds = plain_smem(P->N, a.K, true) <= 160 * 1024;
  for (int s = P->begin; s < P->end; ++s) {  // one launch per sample
    a.ti = (int64_t)s * P->sample;
    int rc = launch_plain_sample(a, P->st);
    if (rc) return rc;
  }
  float *re = P->out.as<float>();
  hipLaunchKernelGGL(d2_to_f, dim3((P->tmax + 255) / 256), dim3(256), 0, P->st, resp, P->tmax, re,
                     re + P->tmax);
  NISE_HIP_CHECK(hipGetLastError());
  return NISE_OK;
}

}  // namespace

// On-the-fly coupling scheme of a Hamiltonian type: 1 TransitionDipole,
// 2 ExtendedDipole (src/read_trajectory.c:180,224), 0 none.
int onfly_scheme(const t_non *non) {
  if (std::strcmp(non->hamiltonian, "TransitionDipole") == 0) return 1;
  if (std::strcmp(non->hamiltonian, "ExtendedDipole") == 0) return 2;
  return 0;
}

// Plan set-up shared by nise_abs_plan_create and nise_abs_plan_create_onfly
// (pos_rec: position records for on-the-fly couplings, else null).
static int plan_create(nise_abs_plan **out, const t_non *non, const float *H_rec, int64_t H_stride,
                       const float *mu_rec, int64_t mu_stride, int64_t nframes, const float *coupling_rec,
                       const float *pos_rec, int64_t pos_stride, const float *box_rec, int where, int kernel) {
  *out = nullptr;
  int rc = ensure_device();
  if (rc) return rc;
  if (!non || !H_rec || !mu_rec) return set_error("null argument"), NISE_ERR_ARG;
  const int onfly = onfly_scheme(non);
  if (onfly && !pos_rec)
    return set_error("%s couplings need position records (nise_abs_plan_create_onfly)", non->hamiltonian),
           NISE_ERR_ARG;
  auto P = new nise_abs_plan();
  std::unique_ptr<nise_abs_plan> guard(P);
  P->non = *non;
  P->N = non->singles, P->tmax = non->tmax1, P->m = non->ts, P->sample = non->sample;
  P->begin = non->begin, P->end = non->end, P->kernel = kernel;
  const int N = P->N;
  if (N < 1 || N > 65535) return set_error("singles=%d out of range", N), NISE_ERR_ARG;
  if (P->tmax < 1 || P->sample < 1 || P->m < 1)
    return set_error("tmax1, sample and Trotter steps must be >= 1"), NISE_ERR_ARG;
  if (non->propagation != 1)
    return set_error("batched absorption supports propagation 1 (Coupling) here; got %d",
                     non->propagation), NISE_ERR_ARG;
  if (P->end == 0) P->end = (int)((nframes - P->tmax - 1) / P->sample + 1);  // src/NISE_subs.c:200,210
  if (P->end <= P->begin) return set_error("no samples (begin=%d end=%d)", P->begin, P->end), NISE_ERR_RANGE;
  P->f_lo = (int64_t)P->begin * P->sample;
  P->f_hi = (int64_t)(P->end - 1) * P->sample + P->tmax;
  if (P->f_hi > nframes)
    return set_error("trajectory has %lld frames, run needs %lld", (long long)nframes,
                     (long long)P->f_hi), NISE_ERR_RANGE;
  const bool ctype = is_coupling_type(non);
  if (ctype && !coupling_rec) return set_error("Coupling type needs coupling_rec"), NISE_ERR_ARG;
  if (H_stride < 1 + ((ctype || onfly) ? N : (int64_t)N * (N + 1) / 2) || mu_stride < 1 + 3 * (int64_t)N ||
      (onfly && pos_stride < 1 + 3 * (int64_t)N))
    return set_error("record stride too small"), NISE_ERR_ARG;
  NISE_HIP_CHECK(hipStreamCreateWithFlags(&P->st, hipStreamNonBlocking));
  for (auto &e : P->ev) NISE_HIP_CHECK(hipEventCreate(&e));

  // device-resident inputs (upload the used frame range when given on host)
  const float *Hd = H_rec, *mud = mu_rec, *Cd = coupling_rec;
  int64_t mu_eff_stride = ctype ? 0 : mu_stride;
  if (where == NISE_HOST) {
    size_t hbytes = (size_t)(P->f_hi - P->f_lo) * H_stride * 4;
    if ((rc = upload(P->H_own, H_rec + P->f_lo * H_stride, hbytes, P->st))) return rc;
    Hd = P->H_own.as<float>() - P->f_lo * H_stride;
    if (ctype) {
      if ((rc = upload(P->mu_own, mu_rec, (size_t)mu_stride * 4, P->st))) return rc;
      mud = P->mu_own.as<float>();
      if ((rc = upload(P->C_own, coupling_rec, (size_t)(1 + (int64_t)N * (N + 1) / 2) * 4, P->st)))
        return rc;
      Cd = P->C_own.as<float>();
    } else {
      if ((rc = upload(P->mu_own, mu_rec + P->f_lo * mu_stride,
                       (size_t)(P->f_hi - P->f_lo) * mu_stride * 4, P->st)))
        return rc;
      mud = P->mu_own.as<float>() - P->f_lo * mu_stride;
    }
  }
  if (onfly) {  // positions (two records per frame for ExtendedDipole) and boxes
    const int64_t per = onfly == 2 ? 2 : 1;
    const float *pd = pos_rec, *bd = box_rec;
    if (where == NISE_HOST) {
      if ((rc = upload(P->pos_own, pos_rec + per * P->f_lo * pos_stride,
                       (size_t)(per * (P->f_hi - P->f_lo) * pos_stride) * 4, P->st)))
        return rc;
      pd = P->pos_own.as<float>() - per * P->f_lo * pos_stride;
      if (box_rec) {
        if ((rc = upload(P->box_own, box_rec + P->f_lo * 4, (size_t)(P->f_hi - P->f_lo) * 16, P->st))) return rc;
        bd = P->box_own.as<float>() - P->f_lo * 4;
      }
    }
    P->src.onfly = onfly, P->src.pos_rec = pd, P->src.pos_stride = pos_stride, P->src.box_rec = bd;
    P->src.dip_rec = mud, P->src.dip_stride = mu_stride;
  }
  P->mu_dev = mud, P->mu_stride = mu_eff_stride;
  P->src.H_rec = Hd, P->src.H_stride = H_stride, P->src.coupling_type = ctype, P->src.C_rec = Cd;
  P->src.shifte = non->shifte, P->src.cut = non->couplingcut, P->src.apply_shift = 1;
  P->src.nshift = non->SingleShiftSites;
  if (P->src.nshift > 0) {
    for (int i = 0; i < P->src.nshift; ++i)
      if (non->SingleShiftSite[i] < 0 || non->SingleShiftSite[i] >= N)
        return set_error("single shift site %d out of range", non->SingleShiftSite[i]), NISE_ERR_ARG;
    if ((rc = upload(P->shift_site, non->SingleShiftSite, P->src.nshift * sizeof(int), P->st))) return rc;
    if ((rc = upload(P->shift_ab, non->SingleShift, 2 * P->src.nshift * sizeof(float), P->st))) return rc;
    P->src.shift_site = P->shift_site.as<int>(), P->src.shift_ab = P->shift_ab.as<float>();
  }
  if ((rc = P->out.alloc(2 * P->tmax * sizeof(float)))) return rc;

  // structure: union of kept couplings over the frames in use
  int64_t npair = (int64_t)N * (N - 1) / 2;
  std::vector<uint8_t> mask(npair);
  if (npair) {
    DevBuf mk;
    if ((rc = mk.alloc(npair))) return rc;
    if ((rc = launch_union_mask(P->src, N, P->f_lo, P->f_hi - P->f_lo, mk.as<uint8_t>(), P->st))) return rc;
    NISE_HIP_CHECK(hipMemcpyAsync(mask.data(), mk.p, npair, hipMemcpyDeviceToHost, P->st));
    NISE_HIP_CHECK(hipStreamSynchronize(P->st));
  }
  build_schedule(N, mask, P->sch_asap);
  if (kernel == 1) {  // the plain per-sample kernel sweeps the layered schedule
    if ((rc = upload(P->ab, P->sch_asap.ab.data(), P->sch_asap.ab.size() * 4, P->st))) return rc;
    if ((rc = upload(P->layer_off, P->sch_asap.layer_off.data(), P->sch_asap.layer_off.size() * 4, P->st)))
      return rc;
    if (plain_smem(N, P->sch_asap.K, true) > 160 * 1024 &&
        (rc = P->ws_cs.alloc((size_t)std::max(P->sch_asap.K, 1) * sizeof(float2))))
      return rc;
    if ((rc = P->resp_plain.alloc(P->tmax * sizeof(double2)))) return rc;
  } else {
    int64_t nfr = P->f_hi - P->f_lo;
    if ((rc = P->diag.alloc((size_t)nfr * N * sizeof(float2)))) return rc;
    if ((rc = P->cs.alloc((size_t)nfr * std::max(P->sch_asap.K, 1) * sizeof(float2)))) return rc;
    // default: the row conveyor with rows matched to the layer width; the
    // slot conveyor's shape (set_shape) stays available
    P->vp = 16, P->spw = 5;
    P->rows = rows_for(N, P->sch_asap.Kreal, P->sch_asap.nlayers());
    build_schedule_rows(N, mask, P->rows, P->sch_rows);
    P->mask = mask;
    if ((rc = build_slots(P))) return rc;
  }
  NISE_HIP_CHECK(hipStreamSynchronize(P->st));
  *out = guard.release();
  return NISE_OK;
}

extern "C" int nise_abs_plan_create(nise_abs_plan **out, const t_non *non, const float *H_rec,
                                    int64_t H_stride, const float *mu_rec, int64_t mu_stride,
                                    int64_t nframes, const float *coupling_rec, int where,
                                    int kernel) {
  return plan_create(out, non, H_rec, H_stride, mu_rec, mu_stride, nframes, coupling_rec, nullptr, 0, nullptr,
                     where, kernel);
}

extern "C" int nise_abs_plan_create_onfly(nise_abs_plan **out, const t_non *non, const float *H_rec,
                                          int64_t H_stride, const float *mu_rec, int64_t mu_stride,
                                          int64_t nframes, const float *pos_rec, int64_t pos_stride,
                                          const float *box_rec, int where, int kernel) {
  if (!non || !onfly_scheme(non))
    return set_error("on-the-fly couplings need HamiltonianType TransitionDipole or ExtendedDipole"), NISE_ERR_ARG;
  if (!pos_rec) return set_error("null position records"), NISE_ERR_ARG;
  return plan_create(out, non, H_rec, H_stride, mu_rec, mu_stride, nframes, nullptr, pos_rec, pos_stride, box_rec,
                     where, kernel);
}

extern "C" int nise_abs_plan_enable_adjoint(nise_abs_plan *P, int every);

extern "C" int nise_abs_plan_set_shape(nise_abs_plan *P, int vp, int samples_per_wg, int block) {
  if (!P || P->kernel != 0) return set_error("no conveyor plan"), NISE_ERR_ARG;
  if (!(vp == 4 || vp == 8 || vp == 16 || vp == 32 || vp == 64) || samples_per_wg < 1 ||
      3 * samples_per_wg > vp ||
      !(block == 64 || block == 256 || block == 512 || block == -2 || block == -4) ||
      (block > 0 && vp > block))
    return set_error("bad shape vp=%d spw=%d block=%d", vp, samples_per_wg, block), NISE_ERR_ARG;
  P->vp = vp, P->spw = samples_per_wg, P->block = block, P->rows = 0;
  if (P->graph_ready) {
    (void)hipGraphExecDestroy(P->exec), (void)hipGraphDestroy(P->graph);
    P->exec = nullptr, P->graph = nullptr, P->graph_ready = false;
  }
  int rc = build_slots(P);
  if (rc) return rc;
  return P->every ? nise_abs_plan_enable_adjoint(P, P->every) : NISE_OK;
}

// Row conveyor with R rows (1, 2, 4, 8 or 16; 0: chosen from the schedule).
extern "C" int nise_abs_plan_set_rows(nise_abs_plan *P, int R) {
  if (!P || P->kernel != 0) return set_error("no conveyor plan"), NISE_ERR_ARG;
  if (R == 0) R = rows_for(P->N, P->sch_asap.Kreal, P->sch_asap.nlayers());
  if (!(R == 1 || R == 2 || R == 4 || R == 8 || R == 16 || R == 32 || R == 64))
    return set_error("rows must be 1, 2, 4, 8, 16, 32 or 64, got %d", R), NISE_ERR_ARG;
  if (rows_smem(P->N, R) > 160 * 1024)
    return set_error("N=%d does not fit the row conveyor with %d rows", P->N, R), NISE_ERR_ARG;
  if (R != P->rows || P->sch_rows.ab.empty()) build_schedule_rows(P->N, P->mask, R, P->sch_rows);
  P->rows = R;
  if (P->graph_ready) {
    (void)hipGraphExecDestroy(P->exec), (void)hipGraphDestroy(P->graph);
    P->exec = nullptr, P->graph = nullptr, P->graph_ready = false;
  }
  int rc = build_slots(P);
  if (rc) return rc;
  return P->every ? nise_abs_plan_enable_adjoint(P, P->every) : NISE_OK;
}

extern "C" int nise_abs_plan_execute(nise_abs_plan *P, float *re_S, float *im_S, int flags);

// Autotuner: times every candidate conveyor configuration on the plan's own
// workload (two event-timed runs each, the faster counts) and keeps the
// fastest.  Candidates: the row conveyor with 1..64 rows (those whose LDS
// image fits) and slot-conveyor shapes.  ms[i] (optional, n entries)
// receives the candidates' times in the order of codes[i]: rows R as R,
// slot shapes as -(1000*vp + 10*spw) - ... see nise_abs_plan_info for the
// chosen one.  Returns the number of candidates timed.
extern "C" int nise_abs_plan_autotune(nise_abs_plan *P, float *ms, int *codes, int n) {
  if (!P || P->kernel != 0) return set_error("no conveyor plan"), -NISE_ERR_ARG;
  struct Cand { int rows, vp, spw, block; };
  std::vector<Cand> cands;
  for (int R : {1, 2, 4, 8, 16, 32, 64})
    if (rows_smem(P->N, R) <= 160 * 1024) cands.push_back({R, 0, 0, 0});
  const Cand slot[] = {{0, 8, 2, 256}, {0, 16, 5, 256}, {0, 8, 2, 64}};
  for (const Cand &c : slot) cands.push_back(c);
  // the plan's configuration on entry, restored if a candidate fails to run
  const Cand orig{P->rows, P->vp, P->spw, P->block};
  auto restore = [&]() {
    (void)(orig.rows ? nise_abs_plan_set_rows(P, orig.rows)
                     : nise_abs_plan_set_shape(P, orig.vp, orig.spw, orig.block));
  };
  float best = 1e30f;
  int bi = -1, k = 0;
  for (size_t i = 0; i < cands.size(); ++i) {
    const Cand &c = cands[i];
    int rc = c.rows ? nise_abs_plan_set_rows(P, c.rows) : nise_abs_plan_set_shape(P, c.vp, c.spw, c.block);
    if (rc) continue;  // shape does not fit this system
    float t = 1e30f;
    for (int r = 0; r < 2; ++r) {
      if ((rc = nise_abs_plan_execute(P, nullptr, nullptr, 1))) {
        restore();
        return -rc;
      }
      t = std::min(t, P->last_ms[1]);
    }
    if (k < n) {
      if (ms) ms[k] = t;
      if (codes) codes[k] = c.rows ? c.rows : -(c.vp * 10000 + c.spw * 1000 + (c.block < 0 ? 900 - c.block : c.block));
    }
    ++k;
    if (t < best) best = t, bi = (int)i;
  }
  if (bi < 0) {
    restore();
    return set_error("autotune: no configuration ran"), -NISE_ERR_ARG;
  }
  const Cand &c = cands[bi];
  int rc = c.rows ? nise_abs_plan_set_rows(P, c.rows) : nise_abs_plan_set_shape(P, c.vp, c.spw, c.block);
  return rc ? -rc : k;
}

// flags bit 0: time the stages with HIP events (no graph)
//       bit 1: replay a captured hipGraph of the stages
extern "C" int nise_abs_plan_execute(nise_abs_plan *P, float *re_S, float *im_S, int flags) {
  if (!P) return set_error("null plan"), NISE_ERR_ARG;
  int rc;
  if (P->kernel == 1) {
    NISE_HIP_CHECK(hipEventRecord(P->ev[0], P->st));
    if ((rc = run_plain(P))) return rc;
    NISE_HIP_CHECK(hipEventRecord(P->ev[1], P->st));
    NISE_HIP_CHECK(hipEventSynchronize(P->ev[1]));
    NISE_HIP_CHECK(hipEventElapsedTime(&P->last_ms[1], P->ev[0], P->ev[1]));
    P->last_ms[0] = P->last_ms[2] = 0;
  } else if (flags & 2) {
    if (!P->graph_ready) {
      NISE_HIP_CHECK(hipStreamBeginCapture(P->st, hipStreamCaptureModeThreadLocal));
      int r1 = run_tables(P), r2 = r1 ? r1 : run_conveyor(P), r3 = r2 ? r2 : run_reduce(P);
      hipError_t ce = hipStreamEndCapture(P->st, &P->graph);
      if (r3) return r3;
      if (ce != hipSuccess) return set_error("graph capture: %s", hipGetErrorString(ce)), NISE_ERR_HIP;
      NISE_HIP_CHECK(hipGraphInstantiate(&P->exec, P->graph, nullptr, nullptr, 0));
      P->graph_ready = true;
    }
    NISE_HIP_CHECK(hipGraphLaunch(P->exec, P->st));
  } else {
    NISE_HIP_CHECK(hipEventRecord(P->ev[0], P->st));
    if ((rc = run_tables(P))) return rc;
    NISE_HIP_CHECK(hipEventRecord(P->ev[1], P->st));
    if ((rc = run_conveyor(P))) return rc;
    NISE_HIP_CHECK(hipEventRecord(P->ev[2], P->st));
    if ((rc = run_reduce(P))) return rc;
    NISE_HIP_CHECK(hipEventRecord(P->ev[3], P->st));
    if (flags & 1) {
      NISE_HIP_CHECK(hipEventSynchronize(P->ev[3]));
      for (int i = 0; i < 3; ++i)
        NISE_HIP_CHECK(hipEventElapsedTime(&P->last_ms[i], P->ev[i], P->ev[i + 1]));
    }
  }
  if (re_S || im_S) {
    std::vector<float> h(2 * P->tmax);
    NISE_HIP_CHECK(hipMemcpyAsync(h.data(), P->out.p, h.size() * 4, hipMemcpyDeviceToHost, P->st));
    NISE_HIP_CHECK(hipStreamSynchronize(P->st));
    if (re_S) std::memcpy(re_S, h.data(), P->tmax * 4);
    if (im_S) std::memcpy(im_S, h.data() + P->tmax, P->tmax * 4);
  }
  return NISE_OK;
}

extern "C" int nise_abs_plan_sync(nise_abs_plan *P) {
  NISE_HIP_CHECK(hipStreamSynchronize(P->st));
  return NISE_OK;
}

// info: N, entries per sweep, layers, workgroups, vp, samples per workgroup,
// first frame, frames, block, steps, rows (0: slot conveyor), rotations
extern "C" int nise_abs_plan_info(const nise_abs_plan *P, int64_t *info, int n) {
  int64_t v[] = {P->N, P->sch.K, P->sch.nlayers(), P->nwg, P->vp, P->spw, P->f_lo, P->f_hi - P->f_lo,
                  P->block, P->nsteps, P->rows, P->sch.Kreal};
  int k = (int)(sizeof(v) / sizeof(v[0]));
  for (int i = 0; i < k && i < n; ++i) info[i] = v[i];
  return k;
}

// ms of the last timed execute: frame tables, conveyor (or plain), reduce;
// then the last adjoint run
extern "C" int nise_abs_plan_last_timing(const nise_abs_plan *P, float *ms, int n) {
  for (int i = 0; i < 4 && i < n; ++i) ms[i] = P->last_ms[i];
  return 4;
}

namespace {
void drop_graph(nise_abs_plan *P) {
  if (P->graph_ready) {
    (void)hipGraphExecDestroy(P->exec), (void)hipGraphDestroy(P->graph);
    P->exec = nullptr, P->graph = nullptr, P->graph_ready = false;
  }
}
}  // namespace

extern "C" int nise_abs_plan_set_params(nise_abs_plan *P, const float *delta, float cscale) {
  if (!P || P->kernel != 0) return set_error("no conveyor plan"), NISE_ERR_ARG;
  if (!std::isfinite(cscale)) return set_error("coupling scale must be finite"), NISE_ERR_ARG;
  int rc;
  if (delta) {
    if ((rc = upload(P->delta, delta, (size_t)P->N * 4, P->st))) return rc;
    P->src.delta = P->delta.as<float>();
  } else {
    P->src.delta = nullptr;
  }
  P->src.cscale = cscale;
  drop_graph(P);  // the captured table launch holds the old parameters
  return NISE_OK;
}

extern "C" int nise_abs_plan_enable_adjoint(nise_abs_plan *P, int every) {
  if (!P || P->kernel != 0) return set_error("no conveyor plan"), NISE_ERR_ARG;
  if (every < 1) return set_error("checkpoint interval must be >= 1"), NISE_ERR_ARG;
  const int T = P->tmax, N = P->N;
  const int64_t nslot = 3 * (int64_t)(P->end - P->begin);
  const int nck = (T - 1) / every + 1 + ((T - 1) % every ? 1 : 0);
  if (nslot * nck > INT32_MAX) return set_error("too many checkpoints"), NISE_ERR_RANGE;
  int rc;
  if ((rc = P->ck.alloc((size_t)nslot * nck * N * sizeof(float2)))) return rc;
  if ((rc = P->jf.alloc((size_t)(P->f_hi - P->f_lo) * std::max(P->sch.K, 1) * 4))) return rc;
  P->every = every, P->nck = nck;
  // backward grouping: consecutive slots, adj_vp per single-wave workgroup
  P->adj_vp = adjoint_vp(P->sch, N, T);
  std::vector<SlotDesc> sl;
  std::vector<int> wb;
  for (int64_t j = 0; j < nslot; ++j) {
    if (j % P->adj_vp == 0) wb.push_back((int)j);
    const int s = (int)(j / 3), x = (int)(j % 3);
    sl.push_back(SlotDesc{(P->begin + s) * P->sample, T, T - 1, x, 0, -1, 0, (int)(j * nck)});
  }
  wb.push_back((int)nslot);
  P->adj_nwg = (int)wb.size() - 1;
  std::vector<int2> stp;
  adjoint_steps(P->sch, P->adj_vp, stp);
  P->adj_nsteps = (int)stp.size();
  if ((rc = upload(P->adj_slots, sl.data(), sl.size() * sizeof(SlotDesc), P->st)) ||
      (rc = upload(P->adj_wg, wb.data(), wb.size() * 4, P->st)) ||
      (rc = upload(P->adj_steps, stp.data(), stp.size() * sizeof(int2), P->st)) ||
      (rc = P->gtab.alloc((size_t)T * sizeof(float2))) ||
      (rc = P->part_d.alloc((size_t)std::max(P->adj_nwg, 1) * N * 8)) ||
      (rc = P->part_c.alloc((size_t)std::max(P->adj_nwg, 1) * 8)) ||
      (rc = P->grad.alloc((size_t)(N + 1) * 8)))
    return rc;
  drop_graph(P);
  return build_slots(P);
}

// Gradient of L with respect to the per-site offsets and the coupling scale
// of the last forward execute, given g = dL/dRe S + i dL/dIm S per step.
extern "C" int nise_abs_plan_adjoint(nise_abs_plan *P, const float *g_re, const float *g_im,
                                     double *grad_delta, double *grad_c) {
  if (!P || !P->every) return set_error("adjoint not enabled on this plan"), NISE_ERR_ARG;
  if (!g_re || !g_im) return set_error("null gradient input"), NISE_ERR_ARG;
  const int T = P->tmax, N = P->N;
  std::vector<float2> gh(T);
  for (int i = 0; i < T; ++i) gh[i] = make_float2(g_re[i], g_im[i]);
  NISE_HIP_CHECK(hipMemcpyAsync(P->gtab.p, gh.data(), T * sizeof(float2), hipMemcpyHostToDevice, P->st));
  AdjArgs a;
  a.N = N, a.K = P->sch.K, a.m = P->m, a.T = T, a.every = P->every, a.nck = P->nck;
  a.fac = trotter_factor(P->non.deltat, 1, P->m);
  a.f_lo = P->f_lo, a.ab = P->ab.as<uint32_t>(), a.cs = P->cs.as<float2>(), a.jf = P->jf.as<float>();
  a.diag = P->diag.as<float2>(), a.mu_rec = P->mu_dev, a.mu_stride = P->mu_stride;
  a.slots = P->adj_slots.as<SlotDesc>(), a.wg_begin = P->adj_wg.as<int>(), a.nwg = P->adj_nwg;
  a.steps = P->adj_steps.as<int2>(), a.nsteps = P->adj_nsteps;
  a.ck = P->ck.as<float2>(), a.g = P->gtab.as<float2>();
  a.part_delta = P->part_d.as<double>(), a.part_c = P->part_c.as<double>();
  NISE_HIP_CHECK(hipEventRecord(P->ev[0], P->st));
  int rc;
  if (P->rows) {  // time-skewed row variant on the row conveyor's tables
    AdjRowsArgs r;
    r.N = N, r.T = T, r.m = P->m, r.nlayers = P->sch.nlayers(), r.every = P->every, r.nck = P->nck;
    r.nvec = 3 * (P->end - P->begin), r.begin = P->begin, r.sample = P->sample;
    r.K = P->sch.K, r.f_lo = P->f_lo, r.fac = a.fac;
    r.steps = P->row_steps.as<uint2>(), r.cs = P->cs.as<float2>(), r.jf = P->jf.as<float>();
    r.diag = P->diag.as<float2>(), r.mu_rec = P->mu_dev, r.mu_stride = P->mu_stride;
    r.ck = P->ck.as<float2>(), r.g = P->gtab.as<float2>();
    const int nwg = (r.nvec + 64 / P->rows - 1) / (64 / P->rows);
    if ((rc = P->part_d.alloc((size_t)std::max(nwg, 1) * N * 8)) ||
        (rc = P->part_c.alloc((size_t)std::max(nwg, 1) * 8)))
      return rc;
    r.part_delta = P->part_d.as<double>(), r.part_c = P->part_c.as<double>();
    rc = launch_adjoint_rows(r, P->rows, P->grad.as<double>(), P->st);
  } else {
    rc = launch_adjoint(a, P->adj_vp, P->grad.as<double>(), P->st);
  }
  if (rc) return rc;
  NISE_HIP_CHECK(hipEventRecord(P->ev[1], P->st));
  std::vector<double> out(N + 1);
  NISE_HIP_CHECK(hipMemcpyAsync(out.data(), P->grad.p, (N + 1) * 8, hipMemcpyDeviceToHost, P->st));
  NISE_HIP_CHECK(hipStreamSynchronize(P->st));
  NISE_HIP_CHECK(hipEventElapsedTime(&P->last_ms[3], P->ev[0], P->ev[1]));
  if (grad_delta) std::memcpy(grad_delta, out.data(), N * 8);
  if (grad_c) *grad_c = out[N];
  return NISE_OK;
}

extern "C" void nise_abs_plan_destroy(nise_abs_plan *P) { delete P; }

extern "C" int nise_absorption_batched(const t_non *non, const float *H_rec, int64_t H_stride,
                                       const float *mu_rec, int64_t mu_stride, int64_t nframes,
                                       const float *coupling_rec, int where, int kernel,
                                       float *re_S, float *im_S) {
  if (non && non->propagation == 0) {  // exact exponential: the DIA conveyor
    nise_system y{non, H_rec, H_stride, mu_rec, mu_stride, nframes, coupling_rec};
    return nise_absorption_multi(&y, 1, where, re_S, im_S, nullptr);
  }
  nise_abs_plan *P = nullptr;
  int rc = nise_abs_plan_create(&P, non, H_rec, H_stride, mu_rec, mu_stride, nframes, coupling_rec,
                                where, kernel);
  if (rc) return rc;
  rc = nise_abs_plan_execute(P, re_S, im_S, 0);
  nise_abs_plan_destroy(P);
  return rc;
}

// ---------------------------------------------------------------------------
// Drop-in singles symbols.  Per host thread: one stream and reusable device
// buffers.  Each call: kept-coupling mask and layer schedule from the frame
// (integer bookkeeping on the host), angles, phases and the sweep on the GPU.
// ---------------------------------------------------------------------------
namespace {

struct SinglesCtx {
  DevBuf H, ab, lay, diag, cs, state, slots, wgb, steps;
  std::vector<float> hbuf;
  std::vector<float2> sbuf;
};

SinglesCtx &ctx() {
  static thread_local SinglesCtx c;
  return c;
}

[[noreturn]] void fatal(const char *where) {
  printf("%s: GPU propagation failed: %s\n", where, get_error());
  fflush(stdout);
  fprintf(stderr, "%s: GPU propagation failed: %s\n", where, get_error());
  exit(1);
}

}  // namespace

namespace nise {

// Propagate nvec vectors (columns of cr/ci, stride N) through one frame H
// (packed triangle, diagonal already shifted) with the Trotter scheme.
int coupling_S_host(int N, const float *H, float *cr, float *ci, int nvec, int m, int sign,
                    float deltat, float cut) {
  int rc = ensure_device();
  if (rc) return rc;
  if (N < 1 || m < 1 || nvec < 1) return set_error("bad arguments"), NISE_ERR_ARG;
  hipStream_t st = default_stream();
  SinglesCtx &c = ctx();
  const int64_t nn2 = (int64_t)N * (N + 1) / 2;
  // frame record with a header word, as the table kernel expects
  c.hbuf.resize(nn2 + 1);
  c.hbuf[0] = 0.f;
  std::memcpy(c.hbuf.data() + 1, H, nn2 * 4);
  std::vector<uint8_t> mask((size_t)N * (N - 1) / 2);
  {
    int64_t p = 0;
    for (int a = 0; a < N; ++a)
      for (int b = a + 1; b < N; ++b, ++p) {
        float J = H[sindex(a, b, N)];
        mask[p] = std::fabs(J) > cut;  // src/propagate.c:853
      }
  }
  Schedule sch;
  build_schedule(N, mask, sch);
  int vp = nvec <= 4 ? 4 : nvec <= 8 ? 8 : nvec <= 16 ? 16 : nvec <= 32 ? 32 : 64;
  while (vp > 4 && (size_t)N * vp * 8 > 150 * 1024) vp >>= 1;
  std::vector<SlotDesc> sl(nvec);
  std::vector<int> wb;
  for (int v = 0; v < nvec; ++v) {
    if (v % vp == 0) wb.push_back(v);
    sl[v] = SlotDesc{0, 0, 1, 0, 1, v, 0, -1};
  }
  wb.push_back(nvec);
  c.sbuf.resize((size_t)nvec * N);
  for (int v = 0; v < nvec; ++v)
    for (int i = 0; i < N; ++i) c.sbuf[(size_t)v * N + i] = make_float2(cr[(size_t)v * N + i], ci[(size_t)v * N + i]);
  if ((rc = upload(c.H, c.hbuf.data(), (nn2 + 1) * 4, st))) return rc;
  if ((rc = upload(c.ab, sch.ab.data(), sch.ab.size() * 4, st))) return rc;
  if ((rc = upload(c.lay, sch.layer_off.data(), sch.layer_off.size() * 4, st))) return rc;
  if ((rc = upload(c.state, c.sbuf.data(), c.sbuf.size() * sizeof(float2), st))) return rc;
  if ((rc = upload(c.slots, sl.data(), sl.size() * sizeof(SlotDesc), st))) return rc;
  if ((rc = upload(c.wgb, wb.data(), wb.size() * sizeof(int), st))) return rc;
  std::vector<int2> stp;
  build_steps(sch, vp, 256, stp);
  if ((rc = upload(c.steps, stp.data(), stp.size() * sizeof(int2), st))) return rc;
  if ((rc = c.diag.alloc(N * sizeof(float2)))) return rc;
  if ((rc = c.cs.alloc(std::max(sch.K, 1) * sizeof(float2)))) return rc;
  FrameSource src;
  src.H_rec = c.H.as<float>(), src.H_stride = nn2 + 1, src.shifte = 0.f, src.cut = cut;
  src.apply_shift = 0;
  float fac = trotter_factor(deltat, sign, m);
  if ((rc = launch_frame_tables(src, N, sch, c.ab.as<uint32_t>(), fac, 0, 1, c.diag.as<float2>(),
                                c.cs.as<float2>(), st)))
    return rc;
  ConveyorArgs a;
  a.N = N, a.K = sch.K, a.nlayers = sch.nlayers(), a.m = m, a.f_lo = 0;
  a.ab = c.ab.as<uint32_t>(), a.layer_off = c.lay.as<int>();
  a.diag = c.diag.as<float2>(), a.cs = c.cs.as<float2>();
  a.mu_rec = c.H.as<float>(), a.mu_stride = 0;
  a.slots = c.slots.as<SlotDesc>(), a.wg_slot_begin = c.wgb.as<int>(), a.nwg = (int)wb.size() - 1;
  a.state = c.state.as<float2>(), a.writeback = 1;
  a.steps = c.steps.as<int2>(), a.nsteps = (int)stp.size();
  if ((rc = launch_conveyor(a, vp, 256, st))) return rc;
  NISE_HIP_CHECK(hipMemcpyAsync(c.sbuf.data(), c.state.p, c.sbuf.size() * sizeof(float2),
                                hipMemcpyDeviceToHost, st));
  NISE_HIP_CHECK(hipStreamSynchronize(st));
  for (int v = 0; v < nvec; ++v)
    for (int i = 0; i < N; ++i) {
      cr[(size_t)v * N + i] = c.sbuf[(size_t)v * N + i].x;
      ci[(size_t)v * N + i] = c.sbuf[(size_t)v * N + i].y;
    }
  return NISE_OK;
}

}  // namespace nise

// src/propagate.c:821-899
extern "C" void propagate_vec_coupling_S(t_non *non, float *Hamiltonian_i, float *cr, float *ci,
                                         int m, int sign) {
  if (coupling_S_host(non->singles, Hamiltonian_i, cr, ci, 1, m, sign, non->deltat,
                      non->couplingcut))
    fatal("propagate_vec_coupling_S");
}

// The sparse scheme's report (src/propagate.c:29-35), printed by the first
// sample's propagation when display is 0.
static void sparse_report(const t_non *non, int elements, int samples, int display) {
  if (samples != non->begin || display != 0) return;
  const int N = non->singles;
  const float scale = non->deltat * kIcm2ifs * kTwoPi / non->ts;
  printf("Sparce matrix efficiency: %f pct.\n", (1 - (1.0 * elements / (N * N))) * 100);
  printf("Pressent tuncation %f.\n", non->thres / (scale * scale));
  printf("Suggested truncation %f.\n", 0.001);
}

// src/propagate.c:20-39
extern "C" void propagate_vector(t_non *non, float *Hamil_i_e, float *vecr, float *veci, int sign,
                                 int samples, int display) {
  if (non->propagation == 1) {
    if (coupling_S_host(non->singles, Hamil_i_e, vecr, veci, 1, non->ts, sign, non->deltat,
                        non->couplingcut))
      fatal("propagate_vector");
    return;
  }
  if (non->propagation == 0 && (non->thres == 0 || non->thres > 1)) {
    if (dia_propagate_host(non, Hamil_i_e, vecr, veci, sign, 1, default_stream()))
      fatal("propagate_vector");
    return;
  }
  if (non->propagation == 0) {  // threshold in (0, 1]: the sparse scheme
    int elements = 0;
    if (dia_s_propagate_host(non, Hamil_i_e, vecr, veci, sign, 1, default_stream(), &elements))
      fatal("propagate_vector");
    sparse_report(non, elements, samples, display);
    return;
  }
  printf("propagate_vector: propagation scheme %d (threshold %g) is not provided by the GPU "
         "library\n", non->propagation, non->thres);
  exit(1);
}

// src/propagate.c:43-52: the N columns of an N x N block, one launch
extern "C" void propagate_matrix(t_non *non, float *Hamil_i_e, float *vecr, float *veci, int sign,
                                 int samples, int display) {
  int N = non->singles;
  if (non->propagation == 1) {
    if (coupling_S_host(N, Hamil_i_e, vecr, veci, N, non->ts, sign, non->deltat, non->couplingcut))
      fatal("propagate_matrix");
    return;
  }
  if (non->propagation == 0 && (non->thres == 0 || non->thres > 1)) {
    if (dia_propagate_host(non, Hamil_i_e, vecr, veci, sign, N, default_stream()))
      fatal("propagate_matrix");
    return;
  }
  if (non->propagation == 0) {  // the sparse scheme, column j reporting with display * j
    int elements = 0;
    if (dia_s_propagate_host(non, Hamil_i_e, vecr, veci, sign, N, default_stream(), &elements))
      fatal("propagate_matrix");
    for (int j = 0; j < N; ++j) sparse_report(non, elements, samples, display * j);
    return;
  }
  printf("propagate_matrix: propagation scheme %d is not provided by the GPU library\n",
         non->propagation);
  exit(1);
}

// Frame tables of one frame for inspection: D phases (N float2) and the
// (cos, sin) of the kept couplings in lexicographic order (count returned
// in *K).  Same code path as the batched plan (FrameSource + table kernel).
extern "C" int nise_frame_tables(const t_non *non, const float *H_rec, int64_t H_stride,
                                 const float *coupling_rec, int64_t pos, int sign, float *diag_out,
                                 float *cs_out, int *K_out) {
  int rc = ensure_device();
  if (rc) return rc;
  const int N = non->singles;
  const bool ctype = is_coupling_type(non);
  hipStream_t st = default_stream();
  DevBuf H, Cb, ab, dg, cs, ss, sab;
  const int64_t nn2 = (int64_t)N * (N + 1) / 2;
  if ((rc = upload(H, H_rec + pos * H_stride, H_stride * 4, st))) return rc;
  FrameSource src;
  src.H_rec = H.as<float>() - pos * H_stride, src.H_stride = H_stride;
  src.coupling_type = ctype;
  if (ctype) {
    if ((rc = upload(Cb, coupling_rec, (1 + nn2) * 4, st))) return rc;
    src.C_rec = Cb.as<float>();
  }
  src.shifte = non->shifte, src.cut = non->couplingcut, src.apply_shift = 1;
  src.nshift = non->SingleShiftSites;
  if (src.nshift > 0) {
    if ((rc = upload(ss, non->SingleShiftSite, src.nshift * sizeof(int), st))) return rc;
    if ((rc = upload(sab, non->SingleShift, 2 * src.nshift * sizeof(float), st))) return rc;
    src.shift_site = ss.as<int>(), src.shift_ab = sab.as<float>();
  }
  std::vector<uint8_t> mask((size_t)N * (N - 1) / 2, 1);
  Schedule sch;
  build_schedule(N, mask, sch);
  // lexicographic order instead of layer order for inspection
  std::vector<uint32_t> lex;
  for (int a = 0; a < N; ++a)
    for (int b = a + 1; b < N; ++b) lex.push_back((uint32_t)a | ((uint32_t)b << 16));
  sch.ab = lex;
  if ((rc = upload(ab, lex.data(), lex.size() * 4, st))) return rc;
  if ((rc = dg.alloc(N * sizeof(float2))) || (rc = cs.alloc((lex.size() + 1) * sizeof(float2))))
    return rc;
  float fac = trotter_factor(non->deltat, sign, non->ts);
  if ((rc = launch_frame_tables(src, N, sch, ab.as<uint32_t>(), fac, pos, 1, dg.as<float2>(),
                                cs.as<float2>(), st)))
    return rc;
  NISE_HIP_CHECK(hipMemcpyAsync(diag_out, dg.p, N * sizeof(float2), hipMemcpyDeviceToHost, st));
  NISE_HIP_CHECK(hipMemcpyAsync(cs_out, cs.p, lex.size() * sizeof(float2), hipMemcpyDeviceToHost, st));
  NISE_HIP_CHECK(hipStreamSynchronize(st));
  *K_out = (int)lex.size();
  return NISE_OK;
}
