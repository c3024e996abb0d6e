// Device-resident solvers: PCG (src/conjgrad.cpp:36-117), split-Bregman
// (src/bregman.cpp:38-134), Nesterov with line search (src/nagopt.cpp:115-209)
// and the MBIR job drivers (src/mbir_cg.cpp, src/mbir_sb.cpp, src/mbir3.cpp).
// Scalars (dots, step sizes, convergence tests) are FP64.
#pragma once
#include <functional>

#include "ops.h"
#include "toeplitz.h"

namespace lam {

// y = A(x) on device volumes, ordered on the context stream
using OpFn = std::function<void(CVol3, Vol3)>;
using LossFn = std::function<double(CVol3)>;

struct CgStats {
    long long iters = 0;
    long long applies = 0;
};

// x <- cgsolver(A, b, x0) ; x may alias x0.
void cg_solve(Ctx &cx, const long long d[3], const OpFn &A, CVol3 b, CVol3 x0, Vol3 x, long long max_iters,
              double tol, CgStats *st);

// Ab(u) = A(u); the split-Bregman system adds mu*laplacian itself.  When
// A_has_lap is set, A already returns A(u) + mu*laplacian(u) (fused form).
void split_bregman(Ctx &cx, const long long d[3], const OpFn &A, bool A_has_lap, CVol3 yT, CVol3 x0, double lambda,
                   double mu, long long outer_max, long long inner_max, double tol, double xtol, Vol3 x,
                   CgStats *st, long long *outer_done);

void nagopt(Ctx &cx, const long long d[3], const OpFn &grad, const LossFn &loss, Vol3 x, long long max_iters,
            double lipschitz, double tol, double xtol, long long max_inner, long long *outer_done,
            long long *grad_calls);

double lipschitz(Ctx &cx, const long long d[3], const OpFn &A, long long max_iters, double tol, uint64_t seed);

// device buffer holding 3 volumes
struct Vol3Buf {
    DevBuf buf;
    long long n = 0;
    void alloc(long long nn) { n = nn; buf.alloc(sizeof(float) * 3 * (size_t)nn); }
    Vol3 v() const { return Vol3{{buf.as<float>(), buf.as<float>() + n, buf.as<float>() + 2 * n}}; }
    CVol3 cv() const { return CVol3{{buf.as<float>(), buf.as<float>() + n, buf.as<float>() + 2 * n}}; }
};

inline CVol3 cv(Vol3 v) { return CVol3{{v.c[0], v.c[1], v.c[2]}}; }

// ---------------------------------------------------------------------------
// Graph form of the solvers.  The PCG scalars (rs, pAp, alpha, beta) and its
// stopping tests live on the device, so iterations are enqueued without any
// host synchronisation; iterations after a stop are no-ops on the device,
// which gives the same iterate as breaking out of the loop.  Every buffer is
// owned by the workspace and fixed for its life, so one captured step can be
// replayed.  Arithmetic is identical to cg_solve / split_bregman.

// Device scalar block of one PCG workspace.
struct CgScalars {
    double sums[4];      // last reduction (3 components)
    double rs_old, alpha, beta;
    long long iters;     // x updates since the last begin
    long long total;     // x updates over the workspace's life
    int done;            // stopping test met (no further updates)
    int pad;
};

struct FastCg {
    Ctx *cx = nullptr;
    long long d[3] = {0, 0, 0};
    long long n = 0;
    Vol3Buf r, z, p, Ap;
    DevBuf scal;     // CgScalars
    DevBuf red;      // reduction partials
    DevBuf pwork;    // preconditioner workspace
    void init(Ctx &c, const long long dims[3]);
    CgScalars *s() const { return scal.as<CgScalars>(); }
};

// x = x0 (when distinct); r = b - A x; z = M r; p = z; rs = z.r
void fast_cg_begin(FastCg &w, const OpFn &A, CVol3 b, CVol3 x0, Vol3 x);
// one PCG iteration (src/conjgrad.cpp:66-107) with device-side stopping tests
void fast_cg_iter(FastCg &w, const OpFn &A, Vol3 x, double tol);
// blocking reads of the device scalars
CgScalars fast_cg_read(FastCg &w);

// Runs `body` (work enqueued on the context stream) eagerly the first time,
// which creates every lazily allocated plan and buffer, then captures it once
// into a hipGraph and replays the graph on later calls.
class GraphStep {
  public:
    ~GraphStep();
    // returns true when this call replayed the captured graph
    bool run(Ctx &cx, bool use_graph, const std::function<void()> &body);

  private:
    int calls_ = 0;
    hipGraph_t graph_ = nullptr;
    hipGraphExec_t exec_ = nullptr;
};

// MBIR1 solve in graph form: x <- PCG(A, b, x0 = 0 in x); returns iterations
long long cg_graph_solve(Ctx &cx, const long long d[3], const OpFn &A, CVol3 b, Vol3 x, long long max_iters,
                         double tol, bool use_graph, CgStats *st);

// How an MBIR job drives its solver: Plain is the host-driven loop of the
// reference (a host read per test), Async keeps the scalars on the device,
// Graph also replays captured iterations (MBIR3's line search decided on the device).
enum class MbirMode { Plain, Async, Graph };

// One reconstruction job end to end on the context's stream (capi.hip).
// transpose120: each component leaves as array::transpose(., {1, 2, 0}) of the
// cropped volume ([rd1][rd2][rd0], the output order of src/recon.cpp:76-78).
void mbir_run(Ctx &cx, const lam_dataset *ds, int nds, const long long rd[3], const lam_recon_params &pr, float *out3,
              lam_recon_stats *stats, MbirMode mode, bool transpose120 = false);
// Upper bound of the device bytes mbir_run allocates for that job.
size_t mbir_footprint(const lam_dataset *ds, int nds, const long long rd[3], const lam_recon_params &pr);

// split_bregman in graph form; A must include mu * laplacian (the fused operator).
// One outer iteration (rhs, PCG with inner_max iterations, shrink, x change) is one step.
void split_bregman_graph(Ctx &cx, const long long d[3], const OpFn &A, CVol3 yT, CVol3 x0, double lambda, double mu,
                         long long outer_max, long long inner_max, double tol, double xtol, Vol3 x, bool use_graph,
                         CgStats *st, long long *outer_done);

// nagopt in graph form (MBIR3).  The line search's accept/reject and the
// momentum and step updates run in a one-thread kernel on the device, so one
// trial (extrapolation, gradient, step, both losses, Armijo test) is enqueued
// without host reads and captured once; trials after the stop are no-ops on
// the iterate.  terms(u, q, s, qpart, rpart) enqueues the data-term pieces of
// loss(u) = sum_g sum_c sqrt(q[3g+c] - 2 s[c] + yTy) into the device arrays
// q (3 * ngrids) and s (3), with qpart / rpart as its reduction scratch.
using LossTermsFn = std::function<void(CVol3 u, double *q, double *s, double *qpart, double *rpart)>;
void nagopt_graph(Ctx &cx, const long long d[3], const OpFn &grad, const LossTermsFn &terms, int ngrids, double yTy,
                  Vol3 x, long long max_iters, double lipschitz, double tol, double xtol, long long max_inner,
                  bool use_graph, long long *outer_done, long long *grad_calls);

}  // namespace lam
