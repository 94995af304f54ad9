// Discrete adjoint of the FD substep: host sequence of the transposed
// column map and right-hand side.
#include "hc_host.h"

#include "hc_adjoint.h"

// Adjoint of substep t_its at the forward state held in the context (the
// substep's input u, v, w, p): the adjoints of its outputs (u', v', w', p'
// and f' = this substep's f) become those of its inputs (u, v, w, p and
// f_old).
int adjoint_substep(hc_ctx *c, int t_its)
{
    const hc_params &P = c->prm;
    StepArgs A = base_args(c);
    if (P.dts == 2) { A.ra = kRkc1[t_its - 1]; A.rb = kRkc2[t_its - 1]; A.rc = kRkc3[t_its - 1]; }
    else { A.ra = 1.5; A.rb = -0.5; A.rc = 1.0; }
    {
        const double *src4[4] = {c->adj[0], c->adj[1], c->adj[2], c->adj[3]};
        double2 *dst4[4] = {c->s[0], c->s[1], c->s[2], c->s[3]};
        if (int e = launch_r2c_n<double>(c, 4, src4, dst4)) return e;
    }
    if (int e = launch_yfft(c, -1, 4, c->s[0], c->s[1], c->s[2], c->s[3])) return e;
    A.s0 = c->s[0]; A.s1 = c->s[1]; A.s2 = c->s[2]; A.s3 = c->s[3];
    dim3 gcol((c->nl + kColWaves - 1) / kColWaves, c->ny);
#define HC_AC(L_) hipLaunchKernelGGL((k_adj_column<L_>), gcol, dim3(64 * kColWaves), col_lds(c), c->stream, A);
    HC_DISPATCH_L(c->Lcol, HC_AC)
#undef HC_AC
    if (int e = launch_yfft(c, +1, 3, c->s[0], c->s[1], c->s[2], nullptr)) return e;
    C2RArgs X;
    std::memset(&X, 0, sizeof X);
    for (int q = 0; q < 3; ++q) { X.src[q] = c->s[q]; X.dst[q] = c->adq[q]; }
    if (int e = launch_c2r<double>(c, 3, X)) return e;
    const int cur = c->adcur, nxt = cur ^ 1;
    AdjIO io;
    io.qu = c->adq[0]; io.qv = c->adq[1]; io.qw = c->adq[2];
    io.fu = c->adf[cur][0]; io.fv = c->adf[cur][1]; io.fw = c->adf[cur][2];
    io.au = c->adj[0]; io.av = c->adj[1]; io.aw = c->adj[2]; io.ap = c->adj[3];
    io.ou = c->adf[nxt][0]; io.ov = c->adf[nxt][1]; io.ow = c->adf[nxt][2];
    for (int q = 0; q < 3; ++q) io.fb[q] = c->adfb[q];
    for (int q = 0; q < 8; ++q) io.face[q] = c->adface[q];
    const dim3 gp(2048), bp(256);
    hipLaunchKernelGGL(k_adj_fbar, gp, bp, 0, c->stream, A, io);
    if (P.cds == 1) {
        hipLaunchKernelGGL((k_adj_faces<1>), gp, bp, 0, c->stream, A, io);
        hipLaunchKernelGGL((k_adj_gather<1>), gp, bp, 0, c->stream, A, io);
    } else {
        hipLaunchKernelGGL((k_adj_faces<2>), gp, bp, 0, c->stream, A, io);
        hipLaunchKernelGGL((k_adj_gather<2>), gp, bp, 0, c->stream, A, io);
    }
    c->adcur = nxt;
    HIPCHK(hipGetLastError());
    return 0;
}
