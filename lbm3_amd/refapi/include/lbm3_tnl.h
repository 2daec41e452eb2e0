// The part of the TNL library (tnl-project.org) that LBM3 case files name
// directly, implemented on HIP for gfx950: the __cuda_callable__ marker,
// device tags, containers with views (device memory for Devices::Cuda, host
// memory for Devices::Host), StaticArray, the Plus / Min / Max reductions,
// Algorithms::reduce and parallelFor, and Timer.  Uses in the reference cases:
// taylorGreenVortexMain.cu:142-167,207-210 (views, reduce, Timer),
// trackerFunctions.h:17-41 (reduce with a lambda reduction).
#pragma once
#include <hip/hip_runtime.h>
#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#ifndef __cuda_callable__
#define __cuda_callable__ __host__ __device__
#endif

namespace TNL {

inline void hipCheck(hipError_t e, const char *what)
{
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

namespace Devices {
struct Cuda {};
struct Host {};
} // namespace Devices

template <class T> __host__ __device__ inline T abs(const T &x) { return x < T(0) ? -x : x; }
template <class T, class U> __host__ __device__ inline auto min(const T &a, const U &b) -> typename std::common_type<T, U>::type
{
    return a < b ? a : b;
}
template <class T, class U> __host__ __device__ inline auto max(const T &a, const U &b) -> typename std::common_type<T, U>::type
{
    return a > b ? a : b;
}

struct Plus {
    template <class T> __host__ __device__ T operator()(const T &a, const T &b) const { return a + b; }
};
struct Min {
    template <class T> __host__ __device__ T operator()(const T &a, const T &b) const { return b < a ? b : a; }
};
struct Max {
    template <class T> __host__ __device__ T operator()(const T &a, const T &b) const { return a < b ? b : a; }
};

class Timer {
public:
    void reset() { total_ = 0.0; running_ = false; }
    void start() { t0_ = std::chrono::steady_clock::now(); running_ = true; }
    void stop()
    {
        if (running_) total_ += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count();
        running_ = false;
    }
    double getRealTime() const
    {
        double t = total_;
        if (running_) t += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count();
        return t;
    }

private:
    std::chrono::steady_clock::time_point t0_{};
    double total_ = 0.0;
    bool running_ = false;
};

namespace Containers {

template <int N, class T> struct StaticArray {
    T data_[N];
    __host__ __device__ StaticArray() {}
    // every element the same value
    template <class U, class = typename std::enable_if<std::is_arithmetic<U>::value>::type>
    __host__ __device__ explicit StaticArray(const U &v)
    {
        for (int i = 0; i < N; ++i) data_[i] = T(v);
    }
    __host__ __device__ T &operator[](int i) { return data_[i]; }
    __host__ __device__ const T &operator[](int i) const { return data_[i]; }
    __host__ __device__ T &operator()(int i) { return data_[i]; }
    __host__ __device__ const T &operator()(int i) const { return data_[i]; }
    __host__ __device__ T &x() { return data_[0]; }
    __host__ __device__ T &y() { return data_[1]; }
    __host__ __device__ T &z() { return data_[2]; }
    __host__ __device__ static constexpr int getSize() { return N; }
    __host__ __device__ void setValue(const T &v)
    {
        for (int i = 0; i < N; ++i) data_[i] = v;
    }
};

// non-owning view of a 1D array (device or host memory, as the owner)
template <class T> struct ArrayView {
    T *data_ = nullptr;
    long long size_ = 0;
    __host__ __device__ T &operator()(long long i) const { return data_[i]; }
    __host__ __device__ T &operator[](long long i) const { return data_[i]; }
    __host__ __device__ long long getSize() const { return size_; }
    __host__ __device__ T *getData() const { return data_; }
};

template <class T> __global__ void k_fill(T *p, long long n, T v)
{
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = v;
}

// 1D array: owning, or (borrow) a view of memory another owner keeps
template <class T, class Device = Devices::Cuda, class Index = long long> class Vector {
    static constexpr bool onDevice = std::is_same<Device, Devices::Cuda>::value;

public:
    using ViewType = ArrayView<T>;
    using ConstViewType = ArrayView<const T>;

    Vector() = default;
    explicit Vector(Index n) { setSize(n); }
    Vector(const Vector &o) { *this = o; }
    Vector(Vector &&o) noexcept { swap(o); }
    template <class D2, class I2> Vector(const Vector<T, D2, I2> &o) { assign(o); }
    ~Vector() { release(); }
    Vector &operator=(const Vector &o)
    {
        if (this != &o) assign(o);
        return *this;
    }
    Vector &operator=(Vector &&o) noexcept
    {
        swap(o);
        return *this;
    }
    template <class D2, class I2> Vector &operator=(const Vector<T, D2, I2> &o)
    {
        assign(o);
        return *this;
    }

    // a view of memory owned elsewhere (the simulation's buffers)
    void borrow(T *p, long long n)
    {
        release();
        p_ = p;
        n_ = n;
        owned_ = false;
    }
    void setSize(Index n)
    {
        if ((long long)n == n_ && owned_) return;
        release();
        n_ = (long long)n;
        if (n_ <= 0) return;
        if (onDevice) hipCheck(hipMalloc((void **)&p_, sizeof(T) * (size_t)n_), "hipMalloc");
        else p_ = static_cast<T *>(::operator new(sizeof(T) * (size_t)n_));
        owned_ = true;
    }
    void resize(Index n) { setSize(n); }
    Index getSize() const { return (Index)n_; }
    T *getData() const { return p_; }
    ViewType getView() { return ViewType{p_, n_}; }
    ConstViewType getView() const { return ConstViewType{p_, n_}; }
    ConstViewType getConstView() const { return ConstViewType{p_, n_}; }
    void setValue(const T &v)
    {
        if (n_ <= 0) return;
        if (onDevice) {
            hipLaunchKernelGGL(k_fill<T>, dim3((unsigned)((n_ + 255) / 256)), dim3(256), 0, nullptr, p_, n_, v);
            hipCheck(hipDeviceSynchronize(), "Vector::setValue");
        } else {
            std::fill(p_, p_ + n_, v);
        }
    }
    T getElement(Index i) const
    {
        if (!onDevice) return p_[i];
        T v;
        hipCheck(hipMemcpy(&v, p_ + i, sizeof(T), hipMemcpyDeviceToHost), "Vector::getElement");
        return v;
    }
    void setElement(Index i, const T &v)
    {
        if (!onDevice) { p_[i] = v; return; }
        hipCheck(hipMemcpy(p_ + i, &v, sizeof(T), hipMemcpyHostToDevice), "Vector::setElement");
    }
    // host containers index directly
    T &operator[](Index i) { return p_[i]; }
    const T &operator[](Index i) const { return p_[i]; }
    T &operator()(Index i) { return p_[i]; }
    const T &operator()(Index i) const { return p_[i]; }

    template <class T2, class D2, class I2> friend class Vector;

private:
    template <class D2, class I2> void assign(const Vector<T, D2, I2> &o)
    {
        setSize((Index)o.n_);
        if (n_ <= 0) return;
        const bool srcDev = std::is_same<D2, Devices::Cuda>::value;
        const hipMemcpyKind k = onDevice ? (srcDev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice)
                                         : (srcDev ? hipMemcpyDeviceToHost : hipMemcpyHostToHost);
        hipCheck(hipMemcpy(p_, o.p_, sizeof(T) * (size_t)n_, k), "Vector copy");
    }
    void release()
    {
        if (p_ && owned_) {
            if (onDevice) (void)hipFree(p_);
            else ::operator delete(p_);
        }
        p_ = nullptr;
        n_ = 0;
        owned_ = true;
    }
    void swap(Vector &o) noexcept
    {
        std::swap(p_, o.p_);
        std::swap(n_, o.n_);
        std::swap(owned_, o.owned_);
    }
    T *p_ = nullptr;
    long long n_ = 0;
    bool owned_ = true;
};

template <class T, class Device = Devices::Cuda, class Index = long long> using Array = Vector<T, Device, Index>;

// 2D host array (rows x columns, row-major), the shape of the tracker histories
template <class T> class HostArray2D {
public:
    void setSizes(long long r, long long c)
    {
        r_ = r;
        c_ = c;
        v_.assign((size_t)(r * c), T());
    }
    template <int D> long long getSize() const { return D == 0 ? r_ : c_; }
    void setValue(const T &x) { std::fill(v_.begin(), v_.end(), x); }
    T &operator()(long long i, long long j) { return v_[(size_t)(i * c_ + j)]; }
    const T &operator()(long long i, long long j) const { return v_[(size_t)(i * c_ + j)]; }
    T getElement(long long i, long long j) const { return (*this)(i, j); }
    void setElement(long long i, long long j, const T &x) { (*this)(i, j) = x; }

private:
    long long r_ = 0, c_ = 0;
    std::vector<T> v_;
};

} // namespace Containers

namespace Algorithms {

template <class F> __global__ void k_parallel_for(long long begin, long long end, F f)
{
    const long long i = begin + (long long)blockIdx.x * 256 + threadIdx.x;
    if (i < end) f((int)i);
}

template <class Device, class Index, class F> void parallelFor(Index begin, Index end, F f)
{
    static_assert(std::is_same<Device, Devices::Cuda>::value, "parallelFor runs on the GPU");
    const long long n = (long long)end - (long long)begin;
    if (n <= 0) return;
    hipLaunchKernelGGL(k_parallel_for<F>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, nullptr, (long long)begin,
                       (long long)end, f);
    hipCheck(hipDeviceSynchronize(), "parallelFor");
}

// one partial per block, combined in block order on the host: the same result
// for the same input on every run
template <class Fetch, class Reduce, class R>
__global__ __launch_bounds__(256) void k_reduce(long long begin, long long end, Fetch fetch, Reduce op, R identity, R *partial)
{
    __shared__ R buf[256];
    R v = identity;
    Fetch f = fetch;
    for (long long i = begin + (long long)blockIdx.x * 256 + threadIdx.x; i < end; i += (long long)gridDim.x * 256)
        v = op(v, (R)f((int)i));
    buf[threadIdx.x] = v;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) buf[threadIdx.x] = op(buf[threadIdx.x], buf[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = buf[0];
}

template <class Device, class Index, class Fetch, class Reduce, class R>
R reduce(Index begin, Index end, Fetch fetch, Reduce op, R identity)
{
    static_assert(std::is_same<Device, Devices::Cuda>::value, "reduce runs on the GPU");
    const long long n = (long long)end - (long long)begin;
    if (n <= 0) return identity;
    const unsigned blocks = (unsigned)std::min<long long>((n + 255) / 256, 4096);
    R *part = nullptr;
    hipCheck(hipMalloc((void **)&part, sizeof(R) * blocks), "reduce scratch");
    hipLaunchKernelGGL((k_reduce<Fetch, Reduce, R>), dim3(blocks), dim3(256), 0, nullptr, (long long)begin, (long long)end,
                       fetch, op, identity, part);
    std::vector<R> host(blocks);
    const hipError_t e = hipMemcpy(host.data(), part, sizeof(R) * blocks, hipMemcpyDeviceToHost);
    (void)hipFree(part);
    hipCheck(e, "reduce");
    R v = identity;
    for (const R &x : host) v = op(v, x);
    return v;
}

} // namespace Algorithms
} // namespace TNL
