// Test harness for include/iemic_gmres.hpp: a restarted, right-preconditioned
// flexible GMRES written only against the GMRES template contract of I-EMIC
// (GMRESSolverDecl.H:4-16: Model::applyMatrix / applyPrecon; Vector copy,
// default construction, assignment, update, norm, dot, scale, the operations
// GMRESSolver.H:108-186 uses), modified Gram-Schmidt and Givens rotations.
// It is instantiated with iemic::GpuModel and
// std::shared_ptr<iemic::DeviceVector>, as the reference's GMRESSolver would
// be, and exported for the Python tests through one C function.
#include <cmath>
#include <cstdio>
#include <memory>
#include <vector>

#include "iemic_gmres.hpp"

template <typename Model, typename VectorPointer>
static int contract_gmres(Model &model, VectorPointer x, VectorPointer b, double tol, int m, int maxit, int &its,
                          double &resid)
{
    using Vector = typename VectorPointer::element_type;
    its = 0;
    double normb = b->norm();
    if (normb == 0.0) normb = 1.0;
    Vector r(*b), tmp(*b), w(*b);
    model.applyMatrix(*x, tmp);
    r.update(-1.0, tmp, 1.0);   // r = b - A x
    double beta = r.norm();
    resid = beta / normb;
    std::vector<Vector> V(m + 1), Z(m);
    std::vector<std::vector<double>> H(m + 1, std::vector<double>(m, 0.0));
    std::vector<double> s(m + 1), cs(m), sn(m), y(m);
    while (resid > tol && its < maxit) {
        r.scale(1.0 / beta);
        V[0] = r;
        std::fill(s.begin(), s.end(), 0.0);
        s[0] = beta;
        int j = 0;
        while (j < m && its < maxit) {
            model.applyPrecon(V[j], tmp);   // z_j = M^{-1} v_j
            Z[j] = tmp;
            model.applyMatrix(tmp, w);      // w = A z_j
            for (int k = 0; k <= j; k++) {
                H[k][j] = w.dot(V[k]);
                w.update(-H[k][j], V[k], 1.0);
            }
            H[j + 1][j] = w.norm();
            if (H[j + 1][j] > 0.0) w.scale(1.0 / H[j + 1][j]);
            V[j + 1] = w;
            for (int k = 0; k < j; k++) {
                const double t = cs[k] * H[k][j] + sn[k] * H[k + 1][j];
                H[k + 1][j] = -sn[k] * H[k][j] + cs[k] * H[k + 1][j];
                H[k][j] = t;
            }
            const double d = std::hypot(H[j][j], H[j + 1][j]);
            cs[j] = d == 0.0 ? 1.0 : H[j][j] / d;
            sn[j] = d == 0.0 ? 0.0 : H[j + 1][j] / d;
            H[j][j] = cs[j] * H[j][j] + sn[j] * H[j + 1][j];
            H[j + 1][j] = 0.0;
            s[j + 1] = -sn[j] * s[j];
            s[j] = cs[j] * s[j];
            j++;
            its++;
            if (std::fabs(s[j]) / normb <= tol) break;
        }
        for (int i = j - 1; i >= 0; i--) {
            double a = s[i];
            for (int k = i + 1; k < j; k++) a -= H[i][k] * y[k];
            y[i] = a / H[i][i];
        }
        for (int i = 0; i < j; i++) x->update(y[i], Z[i], 1.0);
        model.applyMatrix(*x, tmp);   // the true residual decides
        r = *b;
        r.update(-1.0, tmp, 1.0);
        beta = r.norm();
        resid = beta / normb;
    }
    return resid <= tol ? 0 : 1;
}

// x = A^{-1} b on the context's Jacobian (computed) with its preconditioner;
// host arrays of ndim doubles.  Returns 0 (converged), 1 (not converged) or
// the library's error code.
extern "C" int contract_gmres_solve(iemic_ctx *ctx, const double *b_host, double *x_host, double tol, int m,
                                    int maxit, int *iterations, double *resid)
{
    try {
        int n = 0;
        iemic::check(iemic_ndim(ctx, &n), "ndim");
        iemic::GpuModel model(ctx);
        auto b = std::make_shared<iemic::DeviceVector>(n);
        auto x = std::make_shared<iemic::DeviceVector>(n);
        b->upload(b_host);
        int its = 0;
        double res = 0.0;
        const int rc = contract_gmres(model, x, b, tol, m, maxit, its, res);
        x->download(x_host);
        *iterations = its;
        *resid = res;
        return rc;
    } catch (const iemic::Error &e) {
        std::fprintf(stderr, "contract GMRES: %s\n", e.what());
        return e.code;
    }
}
