// Field.hpp -- device-resident field of one box (the engine's counterpart of
// an AMReX MultiFab with a single FArrayBox): SoA, x fastest, owned device
// memory (hipMalloc), with an mfa_field view for the C ABI.  Two orders of
// the components: the MultiFab's planes ([comp][z][y][x]) and the
// row-interleaved order ([z][y][comp][x]), which the step engine streams
// about 5 % faster; the ABI takes either through its strides.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/mfa_lbm.h"

namespace muphfasa {

inline void hip_check(hipError_t e, const char* what)
{
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

inline void mfa_check(int rc, const char* what)
{
    if (rc != MFA_OK) throw std::runtime_error(std::string(what) + ": " + mfa_last_error());
}

inline size_t dtype_size(int dt)
{
    switch (dt) {
    case MFA_F64: return 8;
    case MFA_F32: return 4;
    default: return 2;
    }
}

enum class Layout { Planes, Rows };

class DeviceField {
public:
    DeviceField() = default;
    DeviceField(const int n[3], int ncomp, int ngrow, int dim, int dtype = MFA_F64,
                Layout layout = Layout::Planes)
    {
        define(n, ncomp, ngrow, dim, dtype, layout);
    }
    ~DeviceField() { release(); }
    DeviceField(const DeviceField&) = delete;
    DeviceField& operator=(const DeviceField&) = delete;
    DeviceField(DeviceField&& o) noexcept { *this = std::move(o); }
    DeviceField& operator=(DeviceField&& o) noexcept
    {
        if (this != &o) {
            release();
            m_base = o.m_base;
            m_bytes = o.m_bytes;
            m_view = o.m_view;
            m_dim = o.m_dim;
            o.m_base = nullptr;
            o.m_bytes = 0;
        }
        return *this;
    }

    void define(const int n[3], int ncomp, int ngrow, int dim, int dtype = MFA_F64,
                Layout layout = Layout::Planes)
    {
        release();
        m_dim = dim;
        const long long e0 = n[0] + 2 * ngrow, e1 = n[1] + 2 * ngrow;
        const long long e2 = n[2] + (dim == 3 ? 2 * ngrow : 0);
        const size_t es = dtype_size(dtype);
        m_bytes = (size_t)(e0 * e1 * e2 * ncomp) * es;
        hip_check(hipMalloc(&m_base, m_bytes ? m_bytes : 8), "hipMalloc(field)");
        m_view.n[0] = n[0];
        m_view.n[1] = n[1];
        m_view.n[2] = n[2];
        m_view.ngrow = ngrow;
        if (layout == Layout::Rows) {
            m_view.sq = e0;
            m_view.sy = e0 * ncomp;
            m_view.sz = e0 * ncomp * e1;
        } else {
            m_view.sy = e0;
            m_view.sz = e0 * e1;
            m_view.sq = e0 * e1 * e2;
        }
        m_view.ncomp = ncomp;
        m_view.dtype = dtype;
        const long long off = ngrow + ngrow * m_view.sy + (dim == 3 ? ngrow : 0) * m_view.sz;
        m_view.data = (char*)m_base + off * es;
    }

    void release()
    {
        if (m_base) (void)hipFree(m_base);
        m_base = nullptr;
        m_bytes = 0;
    }

    mfa_field* view() { return &m_view; }
    const mfa_field* view() const { return &m_view; }
    void* base() const { return m_base; }
    size_t bytes() const { return m_bytes; }
    int ncomp() const { return m_view.ncomp; }
    const int* n() const { return m_view.n; }
    long long cells() const { return (long long)m_view.n[0] * m_view.n[1] * m_view.n[2]; }

    // valid cells, component-major [comp][z][y][x] on the host whatever the
    // device order, element type of the field
    void copyValidToHost(void* dst, hipStream_t s = nullptr) const
    {
        const size_t es = dtype_size(m_view.dtype);
        const size_t row = (size_t)m_view.n[0] * es;
        char* d = (char*)dst;
        hip_check(hipStreamSynchronize(s), "sync");
        for (int c = 0; c < m_view.ncomp; ++c)
            for (int z = 0; z < m_view.n[2]; ++z) {
                const char* src = (const char*)m_view.data +
                                  ((long long)c * m_view.sq + (long long)z * m_view.sz) * es;
                hip_check(hipMemcpy2D(d, row, src, (size_t)m_view.sy * es, row, m_view.n[1],
                                      hipMemcpyDeviceToHost),
                          "hipMemcpy2D D2H");
                d += row * m_view.n[1];
            }
    }
    void copyValidFromHost(const void* srcp, hipStream_t s = nullptr)
    {
        const size_t es = dtype_size(m_view.dtype);
        const size_t row = (size_t)m_view.n[0] * es;
        const char* sp = (const char*)srcp;
        hip_check(hipStreamSynchronize(s), "sync");
        for (int c = 0; c < m_view.ncomp; ++c)
            for (int z = 0; z < m_view.n[2]; ++z) {
                char* dst = (char*)m_view.data + ((long long)c * m_view.sq + (long long)z * m_view.sz) * es;
                hip_check(hipMemcpy2D(dst, (size_t)m_view.sy * es, sp, row, row, m_view.n[1],
                                      hipMemcpyHostToDevice),
                          "hipMemcpy2D H2D");
                sp += row * m_view.n[1];
            }
    }

private:
    void* m_base = nullptr;
    size_t m_bytes = 0;
    mfa_field m_view{};
    int m_dim = 3;
};

}  // namespace muphfasa
