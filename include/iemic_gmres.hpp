/* The GMRES template contract of I-EMIC (GMRESSolverDecl.H:4-16) on this
 * library's device state.
 *
 * The reference's GMRESSolver<Model, VectorPointer> asks of its types:
 *   Model::applyMatrix(v, x)   x = A v
 *   Model::applyPrecon(v, x)   x = M^{-1} v
 *   Vector: copy construction, a default construction and assignment (its
 *           basis is a std::vector<Vector>, GMRESSolver.H:130-131),
 *           update(a, A, b)    this = a A + b this,
 *           norm(), dot(other), scale(s)        (GMRESSolver.H:178-186).
 * iemic::DeviceVector is such a Vector over a device array (value
 * semantics: copies are device-to-device), and iemic::GpuModel such a Model
 * over an iemic_ctx whose Jacobian has been computed: applyMatrix is the
 * Jacobian product (Ocean::applyMatrix), applyPrecon the preconditioner of
 * the context (Ocean::applyPrecon).  With them the reference's template is
 * instantiated as GMRESSolver<iemic::GpuModel,
 * std::shared_ptr<iemic::DeviceVector>> without changes.
 *
 * Every operation returns when its result is in place (the reference's
 * GMRES is host-driven: each dot and norm is read on the host).  Errors
 * throw iemic::Error with the library's message.
 */
#pragma once

#include <stdexcept>
#include <string>
#include <utility>

#include "iemic_amd.h"

namespace iemic {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &what) : std::runtime_error(what), code(c) {}
};

inline void check(int rc, const char *where)
{
    if (rc != IEMIC_OK) throw Error(rc, std::string(where) + ": " + iemic_last_error());
}

class DeviceVector {
  public:
    DeviceVector() = default;
    explicit DeviceVector(long long n) : n_(n)
    {
        allocate();
        if (n_ > 0) check(iemic_vec_axpby_dev(n_, 0.0, d_, 0.0, d_), "DeviceVector");   // zero
    }
    DeviceVector(const DeviceVector &o) : n_(o.n_)
    {
        allocate();
        copy_from(o);
    }
    DeviceVector(DeviceVector &&o) noexcept : n_(o.n_), d_(o.d_)
    {
        o.n_ = 0;
        o.d_ = nullptr;
    }
    DeviceVector &operator=(const DeviceVector &o)
    {
        if (this == &o) return *this;
        if (n_ != o.n_) {
            release();
            n_ = o.n_;
            allocate();
        }
        copy_from(o);
        return *this;
    }
    DeviceVector &operator=(DeviceVector &&o) noexcept
    {
        std::swap(n_, o.n_);
        std::swap(d_, o.d_);
        return *this;
    }
    ~DeviceVector() { release(); }

    // this = a A + b this
    void update(double a, const DeviceVector &A, double b)
    {
        same_size(A, "DeviceVector::update");
        check(iemic_vec_axpby_dev(n_, a, A.d_, b, d_), "DeviceVector::update");
    }
    double dot(const DeviceVector &o) const
    {
        same_size(o, "DeviceVector::dot");
        double r = 0.0;
        check(iemic_vec_dot_dev(n_, d_, o.d_, &r), "DeviceVector::dot");
        return r;
    }
    double norm() const
    {
        const double s = dot(*this);
        return s > 0.0 ? __builtin_sqrt(s) : 0.0;
    }
    void scale(double s) { check(iemic_vec_axpby_dev(n_, 0.0, d_, s, d_), "DeviceVector::scale"); }

    // host copies, for setting up a right-hand side and reading a solution
    void upload(const double *h) { check(iemic_memcpy_h2d(d_, h, n_ * 8), "DeviceVector::upload"); }
    void download(double *h) const { check(iemic_memcpy_d2h(h, d_, n_ * 8), "DeviceVector::download"); }

    long long size() const { return n_; }
    double *data() { return d_; }
    const double *data() const { return d_; }

  private:
    long long n_ = 0;
    double *d_ = nullptr;

    void allocate()
    {
        if (n_ > 0) check(iemic_dev_alloc(reinterpret_cast<void **>(&d_), n_ * 8), "DeviceVector");
    }
    void release()
    {
        if (d_) (void)iemic_dev_free(d_);
        d_ = nullptr;
    }
    void copy_from(const DeviceVector &o)
    {
        if (n_ > 0) check(iemic_vec_axpby_dev(n_, 1.0, o.d_, 0.0, d_), "DeviceVector copy");
    }
    void same_size(const DeviceVector &o, const char *where) const
    {
        if (o.n_ != n_) throw Error(IEMIC_EINVAL, std::string(where) + ": vector lengths differ");
    }
};

// The Model of the contract over one context (its Jacobian computed).
class GpuModel {
  public:
    explicit GpuModel(iemic_ctx *ctx) : ctx_(ctx) {}
    // x = J v (Ocean::applyMatrix, Ocean.C:1369-1374)
    void applyMatrix(const DeviceVector &v, DeviceVector &x)
    {
        prepare(v, x);
        check(iemic_apply_matrix_dev(ctx_, v.data(), x.data()), "GpuModel::applyMatrix");
        check(iemic_synchronize(ctx_), "GpuModel::applyMatrix");
    }
    // x = M^{-1} v (Ocean::applyPrecon, Ocean.C:1394-1411)
    void applyPrecon(const DeviceVector &v, DeviceVector &x)
    {
        prepare(v, x);
        check(iemic_apply_precon_dev(ctx_, v.data(), x.data()), "GpuModel::applyPrecon");
        check(iemic_synchronize(ctx_), "GpuModel::applyPrecon");
    }
    iemic_ctx *context() const { return ctx_; }

  private:
    iemic_ctx *ctx_;
    void prepare(const DeviceVector &v, DeviceVector &x)
    {
        int n = 0;
        check(iemic_ndim(ctx_, &n), "GpuModel");
        if (v.size() != n) throw Error(IEMIC_EINVAL, "GpuModel: vector length is not ndim");
        if (x.size() != n) x = DeviceVector(n);
    }
};

}  // namespace iemic
