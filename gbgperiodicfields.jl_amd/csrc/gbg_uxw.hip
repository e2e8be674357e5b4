// Pipelines: the dealiased rotational nonlinear term and the dealiased
// scalar product.
//
// The composed path (GBG_UXW_UNFUSED) runs the standalone operators in the
// order the reference's consumer does (examples/HIST_CPU.jl:229-247):
//   u^t = M u^;  w^ = curl u^t;  E, Z;  u = F^-1 u^t;  w = F^-1 w^;
//   n = u x w;  N^ = F n;  N^ = P N^;  N^ = M N^;  N^(0) = 0.
#include "gbg_internal.hpp"

namespace gbg {

gbg_status pipe_uxw_composed(gbg_plan* p, const void* u_hat, void* n_hat, double kc,
                             hipStream_t s) {
  const size_t cs = (size_t)p->elem * 2;  // complex element bytes
  const size_t rs = (size_t)p->elem;      // real element bytes
  const size_t spec3 = 3 * p->M * cs, phys3 = 3 * p->P * rs;
  gbg_status st = ensure_work(p, 2 * spec3 + 2 * phys3);
  if (st) return st;
  char* w = (char*)p->work;
  void* ut = w;
  void* wt = w + spec3;
  void* u = w + 2 * spec3;
  void* om = w + 2 * spec3 + phys3;
  prof_begin(p, s);
  if ((st = op_convolve_band(p, ut, u_hat, 3, 0.0, kc, 1.0, 0.0, s))) return st;
  if ((st = op_curl(p, wt, ut, 0, s))) return st;
  if ((st = op_inner(p, ut, ut, 3, 0, s))) return st;
  if ((st = op_inner(p, wt, wt, 3, 1, s))) return st;
  prof_mark(p, s);
  if ((st = inv_transform(p, u, ut, ut, 3, s))) return st;
  if ((st = inv_transform(p, om, wt, wt, 3, s))) return st;
  prof_mark(p, s);
  if ((st = op_cross(p, u, u, om, s))) return st;
  prof_mark(p, s);
  if ((st = fwd_transform(p, n_hat, u, 3, s))) return st;
  prof_mark(p, s);
  if ((st = op_helmholtz(p, n_hat, n_hat, 0, 0.0, 0.0, 1.0, 0.0, s))) return st;
  if ((st = op_convolve_band(p, n_hat, n_hat, 3, 0.0, kc, 1.0, 0.0, s))) return st;
  if ((st = op_mean_free(p, n_hat, 3, s))) return st;
  prof_mark(p, s);
  return GBG_OK;
}

gbg_status pipe_uxw(gbg_plan* p, const void* u_hat, void* n_hat, double kc, unsigned flags,
                    hipStream_t s) {
  (void)flags;
  return pipe_uxw_composed(p, u_hat, n_hat, kc, s);
}

gbg_status pipe_dealiased_product(gbg_plan* p, void* p_hat, const void* a_hat, const void* b_hat,
                                  double kc, hipStream_t s) {
  const size_t cs = (size_t)p->elem * 2, rs = (size_t)p->elem;
  const size_t spec1 = p->M * cs, phys1 = p->P * rs;
  gbg_status st = ensure_work(p, 2 * spec1 + 2 * phys1);
  if (st) return st;
  char* w = (char*)p->work;
  void* at = w;
  void* bt = w + spec1;
  void* a = w + 2 * spec1;
  void* b = w + 2 * spec1 + phys1;
  prof_begin(p, s);
  if ((st = op_convolve_band(p, at, a_hat, 1, 0.0, kc, 1.0, 0.0, s))) return st;
  if ((st = op_convolve_band(p, bt, b_hat, 1, 0.0, kc, 1.0, 0.0, s))) return st;
  if ((st = inv_transform(p, a, at, at, 1, s))) return st;
  if ((st = inv_transform(p, b, bt, bt, 1, s))) return st;
  if ((st = op_multiply(p, a, a, b, s))) return st;
  if ((st = fwd_transform(p, p_hat, a, 1, s))) return st;
  if ((st = op_convolve_band(p, p_hat, p_hat, 1, 0.0, kc, 1.0, 0.0, s))) return st;
  prof_mark(p, s);
  return GBG_OK;
}

}  // namespace gbg
