// grasph2/particles.hpp -- C++ mirror of GraSPH2's particle template API,
// backed by the C ABI of grasph2.h.
//
// The same names as the reference so code written against it compiles with
// a change of include:
//   attributes POS MASS VEL ACC BALSARA XVEL DENSITY DENSITY_DT DSTRESS
//     DSTRESS_DT MAXVSIG          src/particles/partice_attributes.h:91-101
//   Particle<Attrs...>           src/particles/Particle.h:36-144
//   HostParticleBuffer<HOST_...> src/particles/HostParticleBuffer.h:47-100
//   DeviceParticleBuffer<DEV_...> src/particles/DeviceParticleBuffer.h:47-110
//   computeDerivatives(pb)       src/computeDerivatives.h:315-328
//   do_for_each_pair_fast<job>(pb) for the jobs calcDensity,
//     calcBalsaraDensityDTDStressDT, calcAcceleration
//                                src/particles/algorithms.h:156, 318-334
//   integrate(pb, notFirstStep)  src/integration.h:219-244
//   getCurrentTimestep()         src/integration.h:246-255
// The compile-time settings of settings/settings.h become the runtime
// grasph::settings() (a grasph_params), read at every call.
//
// Storage differs from the reference where the device kernels want it: the
// tensors DSTRESS / DSTRESS_DT are symmetric and kept as six planes, and the
// extra storage DEV_POS64 / HOST_POS64 holds FP64 master positions (when a
// device buffer has it, integrate() drifts in FP64).  Host-only code: needs
// <hip/hip_runtime_api.h> for device memory and links libgrasph2_amd and
// libamdhip64.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <tuple>
#include <type_traits>
#include <utility>
#include <vector>

#include "../grasph2.h"

// ------------------------------------------------------------------ value types (src/types.h:64-76)
using f1_t = float;
struct f3_t {
    float x = 0, y = 0, z = 0;
};
inline f3_t operator+(f3_t a, f3_t b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline f3_t operator-(f3_t a, f3_t b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline f3_t operator*(f3_t a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline f3_t operator*(float s, f3_t a) { return a * s; }
inline bool operator==(f3_t a, f3_t b) { return a.x == b.x && a.y == b.y && a.z == b.z; }
inline f3_t operator-(f3_t a) { return {-a.x, -a.y, -a.z}; }
inline f3_t operator/(f3_t a, float s) { return {a.x / s, a.y / s, a.z / s}; }
inline f3_t& operator+=(f3_t& a, f3_t b) { return a = a + b; }
inline f3_t& operator-=(f3_t& a, f3_t b) { return a = a - b; }
inline f3_t& operator*=(f3_t& a, float s) { return a = a * s; }
inline f3_t& operator/=(f3_t& a, float s) { return a = a / s; }
inline float dot(f3_t a, f3_t b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline f3_t cross(f3_t a, f3_t b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline float length(f3_t a) { return std::sqrt(dot(a, a)); }
inline f3_t normalize(f3_t a) { return a / length(a); }

// 3x3 matrix, row-major, Mat(v) = v * identity as mpu::Mat<float,3,3>
struct m3_t {
    float m[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    m3_t() = default;
    explicit m3_t(float v) { m[0][0] = m[1][1] = m[2][2] = v; }
    // nine entries, row by row
    m3_t(float a00, float a01, float a02, float a10, float a11, float a12, float a20, float a21, float a22)
        : m{{a00, a01, a02}, {a10, a11, a12}, {a20, a21, a22}}
    {
    }
    float* operator[](int r) { return m[r]; }
    const float* operator[](int r) const { return m[r]; }
    float& operator()(int e) { return m[e / 3][e % 3]; }
    float operator()(int e) const { return m[e / 3][e % 3]; }
};
inline m3_t operator+(const m3_t& a, const m3_t& b)
{
    m3_t r;
    for (int e = 0; e < 9; ++e) r(e) = a(e) + b(e);
    return r;
}
inline m3_t operator*(float s, const m3_t& a)
{
    m3_t r;
    for (int e = 0; e < 9; ++e) r(e) = s * a(e);
    return r;
}
// matrix times column vector: row r of the result is row r of a dotted with v
inline f3_t operator*(const m3_t& a, f3_t v)
{
    return {a[0][0] * v.x + a[0][1] * v.y + a[0][2] * v.z, a[1][0] * v.x + a[1][1] * v.y + a[1][2] * v.z,
            a[2][0] * v.x + a[2][1] * v.y + a[2][2] * v.z};
}

// ------------------------------------------------------------------ attributes
#define GRASPH_ATTRIBUTE(NAME, member, TYPE)                      \
    struct NAME {                                                 \
        using type = TYPE;                                        \
        static constexpr const char* debugName() { return #NAME; } \
        TYPE member{};                                            \
        NAME() = default;                                         \
        explicit NAME(TYPE v) : member(v) {}                      \
        TYPE& getMemberRef() { return member; }                   \
        const TYPE& getMemberRef() const { return member; }       \
    }

GRASPH_ATTRIBUTE(POS, pos, f3_t);
GRASPH_ATTRIBUTE(MASS, mass, f1_t);
GRASPH_ATTRIBUTE(VEL, vel, f3_t);
GRASPH_ATTRIBUTE(ACC, acc, f3_t);
GRASPH_ATTRIBUTE(BALSARA, balsara, f1_t);
GRASPH_ATTRIBUTE(XVEL, xvel, f3_t);
GRASPH_ATTRIBUTE(DENSITY, density, f1_t);
GRASPH_ATTRIBUTE(DENSITY_DT, density_dt, f1_t);
GRASPH_ATTRIBUTE(DSTRESS, dstress, m3_t);
GRASPH_ATTRIBUTE(DSTRESS_DT, dstress_dt, m3_t);
GRASPH_ATTRIBUTE(MAXVSIG, max_vsig, f1_t);
#undef GRASPH_ATTRIBUTE

namespace grasph {
namespace detail {
template <typename T, typename... L> struct contains : std::false_type {};
template <typename T, typename H, typename... L>
struct contains<T, H, L...> : std::conditional_t<std::is_same<T, H>::value, std::true_type, contains<T, L...>> {};
}  // namespace detail
}  // namespace grasph

// Particle<Attrs...>: one particle holding the listed attributes.
// Converting construction / assignment copies the attributes both sides
// have; the others keep their (zero) defaults, as Particle.h:129-143.
template <typename... Args>
class Particle : public Args... {
public:
    using attributes = std::tuple<Args...>;
    static constexpr size_t numAttributes() { return sizeof...(Args); }

    Particle() = default;
    template <typename... T, std::enable_if_t<(sizeof...(T) == sizeof...(Args)) && (sizeof...(T) > 0) &&
                                                  !std::is_same<std::decay_t<std::tuple_element_t<0, std::tuple<T...>>>,
                                                                Particle>::value,
                                              int> = 0>
    explicit Particle(T&&... v) : Args(std::forward<T>(v))... {}

    template <typename... T>
    Particle(const Particle<T...>& b) { assign(b); }

    template <typename... T>
    Particle& operator=(const Particle<T...>& b) { assign(b); return *this; }

    template <typename A> static constexpr bool has() { return grasph::detail::contains<A, Args...>::value; }
    template <typename A> auto& getAttributeRef() { return A::getMemberRef(); }
    template <typename A> const auto& getAttributeRef() const { return A::getMemberRef(); }
    template <typename A> auto getAttribute() const { return A::getMemberRef(); }
    template <typename A, typename V> void setAttribute(const V& v) { A::getMemberRef() = v; }

private:
    template <typename... T>
    void assign(const Particle<T...>& b)
    {
        int t[] = {0, (copy_one<Args>(b), 0)...};
        (void)t;
    }
    template <typename A, typename... T>
    void copy_one(const Particle<T...>& b)
    {
        copy_if<A>(b, std::integral_constant<bool, Particle<T...>::template has<A>()>{});
    }
    template <typename A, typename P> void copy_if(const P& b, std::true_type) { A::getMemberRef() = b.template getAttributeRef<A>(); }
    template <typename A, typename P> void copy_if(const P&, std::false_type) {}
};

// ------------------------------------------------------------------ settings + context
namespace grasph {

// runtime settings (settings/settings.h); defaults from grasph_params_default
inline grasph_params& settings()
{
    static grasph_params p = [] {
        grasph_params q;
        grasph_params_default(&q);
        return q;
    }();
    return p;
}

// process-wide context on device 0 (created on first use, released by
// shutdown(); call that before returning from main so device memory is
// freed while the HIP runtime is still up)
inline grasph_ctx*& context_slot()
{
    static grasph_ctx* c = nullptr;
    return c;
}
inline grasph_ctx* context()
{
    grasph_ctx*& c = context_slot();
    if (!c && grasph_create(0, &c) != 0) {
        c = nullptr;
        throw std::runtime_error(std::string("grasph_create: ") + grasph_last_error(nullptr));
    }
    return c;
}
inline void shutdown()
{
    grasph_ctx*& c = context_slot();
    if (c) grasph_destroy(c);
    c = nullptr;
}

inline void check(int r)
{
    if (r != 0) throw std::runtime_error(std::string("grasph: ") + grasph_last_error(context()));
}

inline void hip_check(hipError_t e, const char* what)
{
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// storage kinds of one attribute group in the C-ABI layout
enum class Store { POSM, POS64, VEC4, SCALAR, SYM6 };

template <Store S> struct StoreInfo;
template <> struct StoreInfo<Store::POSM> { static size_t bytes(size_t n) { return 16 * n; } };
template <> struct StoreInfo<Store::POS64> { static size_t bytes(size_t n) { return 24 * n; } };
template <> struct StoreInfo<Store::VEC4> { static size_t bytes(size_t n) { return 16 * n; } };
template <> struct StoreInfo<Store::SCALAR> { static size_t bytes(size_t n) { return 4 * n; } };
template <> struct StoreInfo<Store::SYM6> { static size_t bytes(size_t n) { return 24 * n; } };

}  // namespace grasph

// Buffer storage tags.  Each names the C-ABI array it occupies and the
// particle attributes it serves.
#define GRASPH_STORAGE(TAG, KIND, FIELD)                               \
    struct HOST_##TAG {                                                \
        static constexpr grasph::Store kind = grasph::Store::KIND;     \
        static constexpr const char* name() { return #FIELD; }         \
    };                                                                 \
    struct DEV_##TAG {                                                 \
        static constexpr grasph::Store kind = grasph::Store::KIND;     \
        static constexpr const char* name() { return #FIELD; }         \
        using host_type = HOST_##TAG;                                  \
    }
GRASPH_STORAGE(POSM, POSM, posm);
GRASPH_STORAGE(POS64, POS64, pos64);
GRASPH_STORAGE(VEL, VEC4, vel);
GRASPH_STORAGE(ACC, VEC4, acc);
GRASPH_STORAGE(XVEL, VEC4, xvel);
GRASPH_STORAGE(BALSARA, SCALAR, balsara);
GRASPH_STORAGE(DENSITY, SCALAR, density);
GRASPH_STORAGE(DENSITY_DT, SCALAR, density_dt);
GRASPH_STORAGE(MAXVSIG, SCALAR, maxvsig);
GRASPH_STORAGE(DSTRESS, SYM6, dstress);
GRASPH_STORAGE(DSTRESS_DT, SYM6, dstress_dt);
#undef GRASPH_STORAGE

namespace grasph {
namespace detail {

inline void** slot(grasph_particles& v, const std::string& name)
{
    if (name == "posm") return (void**)&v.posm;
    if (name == "pos64") return (void**)&v.pos64;
    if (name == "vel") return (void**)&v.vel;
    if (name == "acc") return (void**)&v.acc;
    if (name == "xvel") return (void**)&v.xvel;
    if (name == "density") return (void**)&v.density;
    if (name == "density_dt") return (void**)&v.density_dt;
    if (name == "balsara") return (void**)&v.balsara;
    if (name == "maxvsig") return (void**)&v.maxvsig;
    if (name == "dstress") return (void**)&v.dstress;
    if (name == "dstress_dt") return (void**)&v.dstress_dt;
    throw std::logic_error("unknown storage " + name);
}

inline size_t bytes_of(Store k, size_t n)
{
    switch (k) {
        case Store::POSM: return StoreInfo<Store::POSM>::bytes(n);
        case Store::POS64: return StoreInfo<Store::POS64>::bytes(n);
        case Store::VEC4: return StoreInfo<Store::VEC4>::bytes(n);
        case Store::SCALAR: return StoreInfo<Store::SCALAR>::bytes(n);
        default: return StoreInfo<Store::SYM6>::bytes(n);
    }
}

// one particle <-> the arrays of a host view
template <typename P>
void load_into(P& p, const grasph_particles& v, size_t i)
{
    const size_t n = (size_t)v.n;
    if constexpr (P::template has<POS>()) {
        if (v.pos64) p.pos = {(float)v.pos64[i], (float)v.pos64[n + i], (float)v.pos64[2 * n + i]};
        else if (v.posm) p.pos = {v.posm[4 * i], v.posm[4 * i + 1], v.posm[4 * i + 2]};
    }
    if constexpr (P::template has<MASS>()) if (v.posm) p.mass = v.posm[4 * i + 3];
    auto v3 = [&](const float* a) { return f3_t{a[4 * i], a[4 * i + 1], a[4 * i + 2]}; };
    auto t6 = [&](const float* a) {
        m3_t t;
        t[0][0] = a[i]; t[1][1] = a[n + i]; t[2][2] = a[2 * n + i];
        t[0][1] = t[1][0] = a[3 * n + i];
        t[0][2] = t[2][0] = a[4 * n + i];
        t[1][2] = t[2][1] = a[5 * n + i];
        return t;
    };
    if constexpr (P::template has<VEL>()) if (v.vel) p.vel = v3(v.vel);
    if constexpr (P::template has<ACC>()) if (v.acc) p.acc = v3(v.acc);
    if constexpr (P::template has<XVEL>()) if (v.xvel) p.xvel = v3(v.xvel);
    if constexpr (P::template has<BALSARA>()) if (v.balsara) p.balsara = v.balsara[i];
    if constexpr (P::template has<DENSITY>()) if (v.density) p.density = v.density[i];
    if constexpr (P::template has<DENSITY_DT>()) if (v.density_dt) p.density_dt = v.density_dt[i];
    if constexpr (P::template has<MAXVSIG>()) if (v.maxvsig) p.max_vsig = v.maxvsig[i];
    if constexpr (P::template has<DSTRESS>()) if (v.dstress) p.dstress = t6(v.dstress);
    if constexpr (P::template has<DSTRESS_DT>()) if (v.dstress_dt) p.dstress_dt = t6(v.dstress_dt);
}

// symmetric tensors store their upper triangle
template <typename P>
void store_from(const P& p, grasph_particles& v, size_t i)
{
    const size_t n = (size_t)v.n;
    if constexpr (P::template has<POS>()) {
        if (v.posm) { v.posm[4 * i] = p.pos.x; v.posm[4 * i + 1] = p.pos.y; v.posm[4 * i + 2] = p.pos.z; }
        if (v.pos64) { v.pos64[i] = p.pos.x; v.pos64[n + i] = p.pos.y; v.pos64[2 * n + i] = p.pos.z; }
    }
    if constexpr (P::template has<MASS>()) if (v.posm) v.posm[4 * i + 3] = p.mass;
    auto s3 = [&](float* a, f3_t x) { a[4 * i] = x.x; a[4 * i + 1] = x.y; a[4 * i + 2] = x.z; a[4 * i + 3] = 0.f; };
    auto s6 = [&](float* a, const m3_t& t) {
        a[i] = t[0][0]; a[n + i] = t[1][1]; a[2 * n + i] = t[2][2];
        a[3 * n + i] = t[0][1]; a[4 * n + i] = t[0][2]; a[5 * n + i] = t[1][2];
    };
    if constexpr (P::template has<VEL>()) if (v.vel) s3(v.vel, p.vel);
    if constexpr (P::template has<ACC>()) if (v.acc) s3(v.acc, p.acc);
    if constexpr (P::template has<XVEL>()) if (v.xvel) s3(v.xvel, p.xvel);
    if constexpr (P::template has<BALSARA>()) if (v.balsara) v.balsara[i] = p.balsara;
    if constexpr (P::template has<DENSITY>()) if (v.density) v.density[i] = p.density;
    if constexpr (P::template has<DENSITY_DT>()) if (v.density_dt) v.density_dt[i] = p.density_dt;
    if constexpr (P::template has<MAXVSIG>()) if (v.maxvsig) v.maxvsig[i] = p.max_vsig;
    if constexpr (P::template has<DSTRESS>()) if (v.dstress) s6(v.dstress, p.dstress);
    if constexpr (P::template has<DSTRESS_DT>()) if (v.dstress_dt) s6(v.dstress_dt, p.dstress_dt);
}

}  // namespace detail
}  // namespace grasph

template <typename... Args> class DeviceParticleBuffer;

// Host buffer: one host array per storage tag, in the C-ABI layout.
template <typename... Args>
class HostParticleBuffer {
public:
    using full_particle_type = Particle<POS, MASS, VEL, ACC, BALSARA, XVEL, DENSITY, DENSITY_DT, DSTRESS, DSTRESS_DT, MAXVSIG>;

    HostParticleBuffer() = default;
    explicit HostParticleBuffer(size_t n) { resize(n); }
    template <typename... D>
    HostParticleBuffer& operator=(const DeviceParticleBuffer<D...>& d)
    {
        resize(d.size());
        grasph_particles dv = d.getDeviceReference();
        for (auto& s : storage_) {
            void* src = *grasph::detail::slot(dv, s.name);
            if (src)
                grasph::hip_check(hipMemcpy(s.data.data(), src, s.data.size(), hipMemcpyDeviceToHost), "hipMemcpy D2H");
        }
        return *this;
    }

    size_t size() const { return n_; }
    void initialize()
    {
        for (auto& s : storage_) std::fill(s.data.begin(), s.data.end(), 0);
    }
    // the C-ABI arrays this buffer type holds, with their storage kinds
    static std::vector<std::pair<std::string, grasph::Store>> storage_names()
    {
        return {{Args::name(), Args::kind}...};
    }

    // loadParticle<Attrs...>(i): a particle with the requested attributes
    template <typename... P>
    auto loadParticle(size_t i) const
    {
        using T = std::conditional_t<(sizeof...(P) > 0), Particle<P...>, full_particle_type>;
        T p{};
        grasph::detail::load_into(p, view(), i);
        return p;
    }
    template <typename... P>
    void storeParticle(size_t i, const Particle<P...>& p)
    {
        grasph_particles v = view();
        grasph::detail::store_from(p, v, i);
    }

    grasph_particles view() const
    {
        grasph_particles v;
        std::memset(&v, 0, sizeof v);
        v.n = (int64_t)n_;
        for (auto& s : storage_) *grasph::detail::slot(v, s.name) = (void*)s.data.data();
        return v;
    }

private:
    struct Arr {
        std::string name;
        std::vector<unsigned char> data;
    };
    void resize(size_t n)
    {
        n_ = n;
        storage_.clear();
        int t[] = {0, (storage_.push_back(Arr{Args::name(), std::vector<unsigned char>(grasph::detail::bytes_of(Args::kind, n), 0)}), 0)...};
        (void)t;
    }
    size_t n_ = 0;
    std::vector<Arr> storage_;
};

// Device buffer: owns one device array per storage tag (hipMalloc / hipFree).
template <typename... Args>
class DeviceParticleBuffer {
public:
    explicit DeviceParticleBuffer(size_t n = 0) { allocate(n); }
    ~DeviceParticleBuffer() { release(); }
    DeviceParticleBuffer(const DeviceParticleBuffer&) = delete;
    DeviceParticleBuffer& operator=(const DeviceParticleBuffer&) = delete;
    DeviceParticleBuffer(DeviceParticleBuffer&& o) noexcept : n_(o.n_), arrs_(std::move(o.arrs_)) { o.n_ = 0; }
    DeviceParticleBuffer& operator=(DeviceParticleBuffer&& o) noexcept
    {
        release();
        n_ = o.n_;
        arrs_ = std::move(o.arrs_);
        o.n_ = 0;
        return *this;
    }

    // upload from a host buffer (resizing as needed); the FP64 masters are
    // filled from posm when the host buffer has none
    template <typename... H>
    DeviceParticleBuffer& operator=(const HostParticleBuffer<H...>& h)
    {
        if (h.size() != n_) {
            release();
            allocate(h.size());
        }
        grasph_particles hv = h.view();
        for (auto& a : arrs_) {
            void* src = *grasph::detail::slot(hv, a.name);
            if (src) {
                grasph::hip_check(hipMemcpy(a.ptr, src, a.bytes, hipMemcpyHostToDevice), "hipMemcpy H2D");
            } else if (a.name == "pos64" && hv.posm) {
                std::vector<double> p64(3 * n_);
                for (size_t i = 0; i < n_; ++i)
                    for (int d = 0; d < 3; ++d) p64[d * n_ + i] = hv.posm[4 * i + d];
                grasph::hip_check(hipMemcpy(a.ptr, p64.data(), a.bytes, hipMemcpyHostToDevice), "hipMemcpy H2D");
            } else {
                grasph::hip_check(hipMemset(a.ptr, 0, a.bytes), "hipMemset");
            }
        }
        return *this;
    }

    size_t size() const { return n_; }
    void initialize()
    {
        for (auto& a : arrs_) grasph::hip_check(hipMemset(a.ptr, 0, a.bytes), "hipMemset");
    }
    // the non-owning view the kernels take (DeviceParticleBuffer.h:217-221)
    grasph_particles getDeviceReference() const
    {
        grasph_particles v;
        std::memset(&v, 0, sizeof v);
        v.n = (int64_t)n_;
        for (auto& a : arrs_) *grasph::detail::slot(v, a.name) = a.ptr;
        return v;
    }

private:
    struct Arr {
        std::string name;
        void* ptr;
        size_t bytes;
    };
    void allocate(size_t n)
    {
        n_ = n;
        int t[] = {0, (add(Args::name(), Args::kind), 0)...};
        (void)t;
    }
    void add(const char* name, grasph::Store k)
    {
        Arr a{name, nullptr, grasph::detail::bytes_of(k, n_)};
        grasph::hip_check(hipMalloc(&a.ptr, a.bytes ? a.bytes : 1), "hipMalloc");
        grasph::hip_check(hipMemset(a.ptr, 0, a.bytes), "hipMemset");
        arrs_.push_back(a);
    }
    void release()
    {
        for (auto& a : arrs_) (void)hipFree(a.ptr);
        arrs_.clear();
        n_ = 0;
    }
    size_t n_ = 0;
    std::vector<Arr> arrs_;
};

// ------------------------------------------------------------------ jobs and dispatch
// Tag types standing for the reference's job structs; the work itself runs
// in the HIP library.
struct calcDensity {};
struct calcBalsaraDensityDTDStressDT {};
struct calcAcceleration {};
struct fixedTsLeapfrog {};

template <typename job, size_t blockSize = 256, typename... D>
void do_for_each_pair_fast(DeviceParticleBuffer<D...>& pb)
{
    grasph_particles v = pb.getDeviceReference();
    if constexpr (std::is_same<job, calcDensity>::value)
        grasph::check(grasph_calc_density(grasph::context(), &v, &grasph::settings(), nullptr));
    else if constexpr (std::is_same<job, calcBalsaraDensityDTDStressDT>::value)
        grasph::check(grasph_calc_balsara_density_dt_dstress_dt(grasph::context(), &v, &grasph::settings(), nullptr));
    else if constexpr (std::is_same<job, calcAcceleration>::value)
        grasph::check(grasph_calc_acceleration(grasph::context(), &v, &grasph::settings(), nullptr));
    else
        static_assert(sizeof(job) == 0, "unknown pair job");
}

template <typename job, size_t blockSize = 256, typename... D>
void do_for_each(DeviceParticleBuffer<D...>& pb, bool not_first_step)
{
    static_assert(std::is_same<job, fixedTsLeapfrog>::value, "unknown per-particle job");
    grasph_particles v = pb.getDeviceReference();
    grasph::check(grasph_integrate(grasph::context(), &v, &grasph::settings(), not_first_step ? 1 : 0, nullptr));
}

// src/computeDerivatives.h:315-328
template <typename pbT>
void computeDerivatives(pbT& particleBuffer)
{
    grasph_particles v = particleBuffer.getDeviceReference();
    grasph::check(grasph_compute_derivatives(grasph::context(), &v, &grasph::settings(), nullptr));
}

// src/integration.h:219-244 (fixed-step leapfrog)
template <typename pbT>
void integrate(pbT& particleBuffer, bool notFirstStep)
{
    do_for_each<fixedTsLeapfrog>(particleBuffer, notFirstStep);
}

// src/integration.h:246-255
inline f1_t getCurrentTimestep()
{
    return (f1_t)grasph_get_current_timestep(grasph::context(), &grasph::settings());
}
