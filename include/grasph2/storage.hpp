// grasph2/storage.hpp -- GraSPH2's snapshot text files: the asynchronous
// writer ResultStorageManager, the TextFile particle source and main.cu's
// loadParticlesFromFile.
//
// Format (src/ResultStorageManager.cu:101-186, printTextFile and
// attributePrinter): one line per particle; every value of every attribute
// of the stored particle type, in attribute order, is printed with
// std::fixed and digits10 + 1 decimals of f1_t (7 for float) and followed by
// a tab; vectors print their components x y z, matrices their 9 entries in
// row-major order.  The attribute order follows the storage tags of the
// buffer: HOST_POSM gives POS then MASS (settings/outputSettings.h:27 stores
// HostParticleBuffer<HOST_POSM, HOST_VEL, HOST_DENSITY>, i.e. x y z m vx vy
// vz rho).  File name: directory + prefix + start time ("%Y-%m-%d_%H-%M") +
// "_" + time (fixed, same decimals) + ".tsv" (:161-163).
//
// Reading (src/initialConditions/particleSources/TextFile.h:69-165):
// non-blank lines are particles; each value is the text up to the next
// separator, converted with std::stod; a line that runs out of values is an
// error.  The particle type chosen by the caller fixes the columns.
//
// DEV_POS64 / HOST_POS64 (FP64 master positions) have no column: they are
// not part of the reference's particle attributes.
#ifndef GRASPH2_STORAGE_HPP
#define GRASPH2_STORAGE_HPP

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <ctime>
#include <fstream>
#include <iomanip>
#include <limits>
#include <memory>
#include <mutex>
#include <queue>
#include <sstream>
#include <stdexcept>
#include <string>
#include <thread>
#include <utility>
#include <vector>

#include "initial_conditions.hpp"
#include "particles.hpp"

namespace grasph {
namespace detail {

inline void put(std::ostream& s, f1_t v) { s << v << "\t"; }
inline void put(std::ostream& s, const f3_t& v) { s << v.x << "\t" << v.y << "\t" << v.z << "\t"; }
inline void put(std::ostream& s, const m3_t& v)
{
    for (int i = 0; i < 9; ++i) s << v(i) << "\t";
}

// the attributes one storage tag contributes to a snapshot line, in order
template <typename Tag> struct text_columns;
template <> struct text_columns<HOST_POSM> { using type = Particle<POS, MASS>; };
template <> struct text_columns<HOST_POS64> { using type = Particle<>; };
template <> struct text_columns<HOST_VEL> { using type = Particle<VEL>; };
template <> struct text_columns<HOST_ACC> { using type = Particle<ACC>; };
template <> struct text_columns<HOST_XVEL> { using type = Particle<XVEL>; };
template <> struct text_columns<HOST_BALSARA> { using type = Particle<BALSARA>; };
template <> struct text_columns<HOST_DENSITY> { using type = Particle<DENSITY>; };
template <> struct text_columns<HOST_DENSITY_DT> { using type = Particle<DENSITY_DT>; };
template <> struct text_columns<HOST_MAXVSIG> { using type = Particle<MAXVSIG>; };
template <> struct text_columns<HOST_DSTRESS> { using type = Particle<DSTRESS>; };
template <> struct text_columns<HOST_DSTRESS_DT> { using type = Particle<DSTRESS_DT>; };

template <typename... A> struct put_all;
template <> struct put_all<> {
    template <typename P> static void run(std::ostream&, const P&) {}
};
template <typename A, typename... R> struct put_all<A, R...> {
    template <typename P> static void run(std::ostream& s, const P& p)
    {
        put(s, p.template getAttributeRef<A>());
        put_all<R...>::run(s, p);
    }
};

template <typename Buf, typename... A> void put_cols(std::ostream& s, const Buf& b, size_t i, Particle<A...>*)
{
    if constexpr (sizeof...(A) > 0) put_all<A...>::run(s, b.template loadParticle<A...>(i));
}
template <typename Tag, typename Buf> void put_tag(std::ostream& s, const Buf& b, size_t i)
{
    using Cols = typename text_columns<Tag>::type;
    put_cols(s, b, i, (Cols*)nullptr);
}

inline void get(std::istream& s, char sep, f1_t& v)
{
    std::string tok;
    std::getline(s, tok, sep);
    v = (f1_t)std::stod(tok);  // throws std::invalid_argument on a missing value, as the reference does
}
inline void get(std::istream& s, char sep, f3_t& v)
{
    get(s, sep, v.x);
    get(s, sep, v.y);
    get(s, sep, v.z);
}
inline void get(std::istream& s, char sep, m3_t& v)
{
    for (int i = 0; i < 9; ++i) get(s, sep, v(i));
}

template <typename... A> struct get_all;
template <> struct get_all<> {
    template <typename P> static void run(std::istream&, char, P&) {}
};
template <typename A, typename... R> struct get_all<A, R...> {
    template <typename P> static void run(std::istream& s, char sep, P& p)
    {
        get(s, sep, p.template getAttributeRef<A>());
        get_all<R...>::run(s, sep, p);
    }
};
template <typename... A> void get_particle(std::istream& s, char sep, Particle<A...>& p) { get_all<A...>::run(s, sep, p); }

inline std::string timestamp_now()
{
    const std::time_t t = std::time(nullptr);
    std::tm tm{};
    localtime_r(&t, &tm);
    char b[64];
    std::strftime(b, sizeof b, "%Y-%m-%d_%H-%M", &tm);
    return b;
}

}  // namespace detail

// printTextFile (src/ResultStorageManager.cu:159-186) for one buffer
template <typename... Tags>
void writeTextSnapshot(std::ostream& file, const HostParticleBuffer<Tags...>& data)
{
    file << std::fixed << std::setprecision(std::numeric_limits<f1_t>::digits10 + 1);
    for (size_t i = 0; i < data.size(); ++i) {
        int t[] = {0, (detail::put_tag<Tags>(file, data, i), 0)...};
        (void)t;
        file << "\n";
    }
}

template <typename... Tags>
void writeTextSnapshot(const std::string& filename, const HostParticleBuffer<Tags...>& data)
{
    std::ofstream file(filename);
    if (!file.is_open()) throw std::runtime_error("could not open output file " + filename);
    writeTextSnapshot(file, data);
    if (file.fail()) throw std::runtime_error("could not write output file " + filename);
}

// ResultStorageManager (src/ResultStorageManager.h:21-186): at most maxJobs
// snapshots wait before printToFile blocks until the queue drains; one
// worker thread writes them.  A device buffer is duplicated on the device
// (src/ResultStorageManager.h:124-160): the arrays the snapshot type stores
// are copied device-to-device on the null stream, where the library's
// entry points are enqueued by the C++ layer, so the copy holds the state
// after the work queued before printToFile and the simulation may go on at
// once; the worker waits for that copy, downloads it on a stream of its own
// and frees it.  Without device memory for the copy the state is
// downloaded on the caller's thread instead (:148-158).
template <typename HostDiscPT = HostParticleBuffer<HOST_POSM, HOST_VEL, HOST_DENSITY>>
class ResultStorageManager {
public:
    ResultStorageManager(std::string directory, std::string prefix, int maxJobs)
        : dir_(std::move(directory)), prefix_(std::move(prefix)), start_(detail::timestamp_now()), max_(maxJobs)
    {
        if (max_ <= 1) throw std::invalid_argument("ResultStorageManager needs maxJobs > 1");
        worker_ = std::thread([this] { work(); });
    }
    ~ResultStorageManager()
    {
        {
            std::lock_guard<std::mutex> l(m_);
            stop_ = true;
        }
        cv_.notify_all();
        worker_.join();
    }
    ResultStorageManager(const ResultStorageManager&) = delete;
    ResultStorageManager& operator=(const ResultStorageManager&) = delete;

    template <typename... D> void printToFile(const DeviceParticleBuffer<D...>& particles, f1_t time)
    {
        Job j;
        j.time = time;
        j.dev.reset(new DeviceCopy);
        if (!j.dev->make(particles.getDeviceReference())) {
            j.dev.reset();
            j.host.reset(new HostDiscPT());
            *j.host = particles;  // no device memory for a copy: download now
        }
        enqueue(std::move(j));
    }
    template <typename... H> void printToFile(const HostParticleBuffer<H...>& particles, f1_t time)
    {
        Job j;
        j.time = time;
        j.host.reset(new HostDiscPT(particles.size()));
        for (size_t i = 0; i < particles.size(); ++i) j.host->storeParticle(i, particles.loadParticle(i));
        enqueue(std::move(j));
    }
    std::string getStartTime() const { return start_; }
    // the file a snapshot at `time` goes to (:161-163)
    std::string filename(f1_t time) const
    {
        std::ostringstream f;
        f << dir_ << prefix_ << start_ << "_" << std::fixed << std::setprecision(std::numeric_limits<f1_t>::digits10 + 1)
          << time << ".tsv";
        return f.str();
    }
    // block until every queued snapshot is on disk; rethrows a write error
    void flush()
    {
        std::unique_lock<std::mutex> l(m_);
        done_.wait(l, [this] { return q_.empty() && !busy_; });
        if (!err_.empty()) {
            std::string e;
            std::swap(e, err_);
            throw std::runtime_error(e);
        }
    }

private:
    // a device-side duplicate of the arrays HostDiscPT stores
    struct DeviceCopy {
        int device = 0;
        int64_t n = 0;
        std::vector<std::pair<std::string, void*>> arrays;
        hipEvent_t done = nullptr;  // recorded after the copies on the null stream
        ~DeviceCopy()
        {
            (void)hipSetDevice(device);
            for (auto& a : arrays) (void)hipFree(a.second);
            if (done) (void)hipEventDestroy(done);
        }
        bool make(grasph_particles src)
        {
            n = src.n;
            if (hipGetDevice(&device) != hipSuccess) return fail();
            for (const auto& t : HostDiscPT::storage_names()) {
                void* from = *detail::slot(src, t.first);
                if (!from) continue;  // the source lacks it: the column stays zero
                const size_t bytes = detail::bytes_of(t.second, (size_t)n);
                void* to = nullptr;
                if (hipMalloc(&to, bytes ? bytes : 1) != hipSuccess) return fail();
                arrays.emplace_back(t.first, to);
                if (hipMemcpyAsync(to, from, bytes, hipMemcpyDeviceToDevice, nullptr) != hipSuccess) return fail();
            }
            if (hipEventCreateWithFlags(&done, hipEventDisableTiming) != hipSuccess) return fail();
            if (hipEventRecord(done, nullptr) != hipSuccess) return fail();
            return true;
        }
        bool fail()
        {
            (void)hipGetLastError();
            return false;
        }
        // on the worker thread: wait for the copies, download on a stream of its own
        std::unique_ptr<HostDiscPT> download() const
        {
            hip_check(hipSetDevice(device), "hipSetDevice");
            std::unique_ptr<HostDiscPT> h(new HostDiscPT((size_t)n));
            hip_check(hipEventSynchronize(done), "hipEventSynchronize");
            hipStream_t s = nullptr;
            hip_check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreateWithFlags");
            grasph_particles hv = h->view();
            hipError_t e = hipSuccess;
            for (const auto& a : arrays) {
                void* to = *detail::slot(hv, a.first);
                const size_t bytes = detail::bytes_of(kind_of(a.first), (size_t)n);
                if (e == hipSuccess) e = hipMemcpyAsync(to, a.second, bytes, hipMemcpyDeviceToHost, s);
            }
            if (e == hipSuccess) e = hipStreamSynchronize(s);
            (void)hipStreamDestroy(s);
            hip_check(e, "snapshot download");
            return h;
        }
        static Store kind_of(const std::string& name)
        {
            for (const auto& t : HostDiscPT::storage_names())
                if (t.first == name) return t.second;
            throw std::logic_error("unknown storage " + name);
        }
    };
    struct Job {
        std::unique_ptr<HostDiscPT> host;
        std::unique_ptr<DeviceCopy> dev;
        f1_t time = 0;
    };

    void enqueue(Job j)
    {
        std::unique_lock<std::mutex> l(m_);
        done_.wait(l, [this] { return (int)q_.size() < max_; });  // queue full: wait (:113-124)
        q_.push(std::move(j));
        cv_.notify_one();
    }
    void work()
    {
        std::unique_lock<std::mutex> l(m_);
        for (;;) {
            cv_.wait(l, [this] { return stop_ || !q_.empty(); });
            if (q_.empty() && stop_) return;
            auto job = std::move(q_.front());
            q_.pop();
            busy_ = true;
            l.unlock();
            std::string e;
            try {
                if (job.dev) {
                    job.host = job.dev->download();
                    job.dev.reset();
                }
                writeTextSnapshot(filename(job.time), *job.host);
            } catch (const std::exception& x) {
                e = x.what();
            }
            l.lock();
            busy_ = false;
            if (!e.empty() && err_.empty()) err_ = e;
            done_.notify_all();
        }
    }

    const std::string dir_, prefix_, start_;
    const int max_;
    std::queue<Job> q_;
    std::mutex m_;
    std::condition_variable cv_, done_;
    bool stop_ = false, busy_ = false;
    std::string err_;
    std::thread worker_;
};

}  // namespace grasph

namespace ps {

// TextFile<ParticleToRead> (src/initialConditions/particleSources/TextFile.h):
// a particle source over the lines of a snapshot file; particle id is line
// id parsed into ParticleToRead, then the source's modifiers apply.
template <typename ParticleToRead>
class TextFile : public ParticleSource<ParticleToRead, TextFile<ParticleToRead>> {
public:
    explicit TextFile(const std::string& filename, char seperator = '\t') : sep_(seperator)
    {
        std::ifstream file(filename);
        if (!file.is_open()) throw std::runtime_error("Error opening file " + filename);
        std::string line;
        while (std::getline(file, line)) {
            bool blank = true;
            for (char c : line) blank = blank && std::isspace((unsigned char)c);
            if (!blank) lines_.push_back(line);
        }
        this->m_numberOfParticles = lines_.size();
    }
    size_t size() const { return lines_.size(); }
    // line id parsed, without the modifiers
    ParticleToRead particle(size_t id) const
    {
        ParticleToRead p{};
        std::istringstream line(lines_[id]);
        try {
            grasph::detail::get_particle(line, sep_, p);
        } catch (const std::exception&) {
            throw std::runtime_error("Error parsing file at line " + std::to_string(id));
        }
        return p;
    }
    // all particles into a host buffer (what InitGenerator::generate() yields for this source alone)
    template <typename... Tags> void load(HostParticleBuffer<Tags...>& out)
    {
        out = HostParticleBuffer<Tags...>(size());
        for (size_t i = 0; i < size(); ++i) (*this)(out, i, i);
    }

private:
    ParticleToRead generateParticle(size_t id) override { return particle(id); }

    std::vector<std::string> lines_;
    char sep_;
};

}  // namespace ps

// Adds the particles of a snapshot file to a generator, by file ending
// (src/main.cu:43-67): .tsv tab-separated, .csv comma-separated, anything
// else with textfileSeperator.  HDF5 files (.h5, .hdf5) need HighFive,
// which this build does not carry: they raise an error.
template <typename LoadParticleType, typename GeneratorType>
void loadParticlesFromFile(const std::string& filename, GeneratorType& generator, char textfileSeperator = '\t')
{
    const auto dot = filename.find_last_of('.');
    const std::string ending = dot == std::string::npos ? std::string() : filename.substr(dot + 1);
    if (ending == "h5" || ending == "hdf5")
        throw std::runtime_error("HDF5 input is not supported by this build: " + filename);
    const char sep = ending == "tsv" ? '\t' : ending == "csv" ? ',' : textfileSeperator;
    generator.addParticles(ps::TextFile<LoadParticleType>(filename, sep));
}

#endif  // GRASPH2_STORAGE_HPP
