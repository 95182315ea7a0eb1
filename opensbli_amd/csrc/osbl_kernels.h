// osbl_kernels.h -- launch interface of the gfx950 kernels (host side).
#pragma once

#include "osbl_device.h"

namespace osbl {

void launch_halo_periodic(double* Q, int nv, const Grid& g, hipStream_t s);
void launch_extrap0(double* Q, int nv, const Grid& g, int d, int side, hipStream_t s);
void launch_prims(int nd, const double* Q, double* P, const Grid& g, const Phys& ph, hipStream_t s);
void launch_sweep(int nd, int d, int conv, int avg, const double* Q, double* R, const Grid& g, const Phys& ph, int seg,
                  const double* met, const double* theta, bool acc, hipStream_t s);
void launch_sweep_z_rk(int conv, int avg, int mode, const double* Q, double* R, double* Qo, double* Qn, const Grid& g,
                       const Phys& ph, int seg, const double* theta, double cA, double cB, int first, hipStream_t s);
void launch_sensor(int nd, const double* Q, double* theta, const Grid& g, const Phys& ph, const Bnd& b, hipStream_t s);
void launch_update(int nd, int conv, bool visc, int mode, const double* Q, const double* P, double* R, double* Qo,
                   double* Qn, const Grid& g, const Phys& ph, double cA, double cB, int first, hipStream_t s);
void launch_visc_update3d(int mode, const double* Q, double* R, double* Qo, double* Qn, const Grid& g, const Phys& ph,
                          double cA, double cB, int first, int sz, hipStream_t s);
void launch_conv3d(int conv, int mode, const double* Q, double* R, double* Qo, double* Qn, const Grid& g,
                   const Phys& ph, double cA, double cB, int first, int sz, hipStream_t s);
// z segments of the tiled viscous and central-convective kernels for a grid,
// and the convective kernel's block count at segment sz
int visc3d_segment(const Grid& g);
int conv3d_segment(const Grid& g);
long long conv3d_blocks(const Grid& g, int sz);
void launch_update_gen(int nd, int conv, bool visc, int mode, const double* Q, const double* P, double* R,
                       double* Qo, double* Qn, const Grid& g, const Phys& ph, const Bnd& b, double cA, double cB,
                       int first, int suth, hipStream_t s);
void launch_visc_gen3d(const double* Q, const double* P, double* R, const Grid& g, const Phys& ph, const Bnd& b,
                       int suth, int sz, hipStream_t s);
void launch_body_force3d(const double* Q, double* R, const Grid& g, const Phys& ph, hipStream_t s);
void launch_prims_slab(const double* Q, double* P, const Grid& g, const Phys& ph, int suth, hipStream_t s);
void launch_prims_mu(int nd, const double* Q, double* P, const Grid& g, const Phys& ph, hipStream_t s);
void launch_bc_face(int nd, double* Q, double* Q2, const Grid& g, const Phys& ph, int d, int side, int type, int nh,
                    double twall, const double* plane, hipStream_t s);
void launch_periodic_dir(double* Q, int nv, const Grid& g, int d, hipStream_t s);
void launch_fill_face_halo(double* Q, int nv, const Grid& g, int d, int side, const double* plane, hipStream_t s);
void launch_pack(const double* F, double* out, const Grid& g, hipStream_t s);
void launch_count_nan(const double* F, int v0, int v1, unsigned long long* count, const Grid& g, hipStream_t s);
void launch_unpack(const double* in, double* F, const Grid& g, hipStream_t s);
long long halo_points(const Grid& g);
void launch_rhs_vjp(int nd, const double* Q, const double* P, const double* M, double* out, const Grid& g,
                    const Phys& ph, hipStream_t s);
void launch_adj_stage(int nd, const double* Qs, const double* A, const double* B, double alpha, double beta, double c,
                      double* out1, double* out2, const Grid& g, const Phys& ph, hipStream_t s);
void launch_pack_all(const double* F, double* out, const Grid& g, int nv, hipStream_t s);
void launch_adj_stage_char(int nd, int avg, const double* Qs, const double* A, const double* B, double alpha,
                           double beta, double c, double* out1, double* out2, const Grid& g, const Phys& ph, bool acc,
                           hipStream_t s, double* vy = nullptr);
int visc_vjp_fields(int nd);
void launch_unpack_prims(int nd, const double* q, double* Qpad, double* P, const Grid& g, const Phys& ph,
                         hipStream_t s);
void launch_visc_vjp(int nd, const double* q, const double* A, const double* B, double alpha, double beta, double* P,
                     double* Mpad, double* W, double* out, const Grid& g, const Phys& ph, hipStream_t s);
void launch_weno_filter_apply(int nd, double* Q, const double* R, double* theta, const Grid& g, const Phys& ph,
                              bool ducros, bool mach, double control, int walls, hipStream_t s);
void launch_sfd(int nd, double* Q, double* F, const Grid& g, double c, double E, hipStream_t s);
void launch_binomial(int nd, const double* Q, double* Qn, const Grid& g, const double* w, int N, double sigma,
                     const unsigned char* mask, hipStream_t s);
void launch_lincomb(double* out, double a, const double* x, double b, const double* y, long long n, hipStream_t s);

void launch_curv_direction(int nd, int d, int conv, int avg, const double* Q, const double* M, double* PF, double* F,
                           double* R, const Grid& g, const Phys& ph, double delta, hipStream_t s);
void launch_curv_metric(int nd, const double* X, double* M, const Grid& g, const Phys& ph, const int e[3],
                        hipStream_t s);

// record the thread-local message osbl_last_error() returns; returns code
int set_error(int code, const char* msg);

}  // namespace osbl
