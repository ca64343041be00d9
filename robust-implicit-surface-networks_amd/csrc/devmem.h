// Per-context device memory cache: allocations are returned to size-keyed
// free lists and reused, so repeated runs (and the serving loop) do not call
// hipMalloc in steady state.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

namespace rin {

struct HipError : std::runtime_error {
    hipError_t code;
    HipError(hipError_t c, const std::string& what) : std::runtime_error(what), code(c) {}
};

#define RIN_HIP(expr)                                                                                  \
    do {                                                                                               \
        hipError_t _e = (expr);                                                                        \
        if (_e != hipSuccess)                                                                          \
            throw ::rin::HipError(_e, std::string(#expr) + ": " + hipGetErrorString(_e));              \
    } while (0)

class DevMem {
  public:
    ~DevMem() { release_all(); }
    void* alloc(size_t bytes) {
        bytes = round(bytes);
        std::lock_guard<std::mutex> g(mu_);
        auto it = free_.lower_bound(bytes);
        if (it != free_.end() && it->first <= bytes * 2 && !it->second.empty()) {
            void* p = it->second.back();
            it->second.pop_back();
            size_[p] = it->first;
            live_ += it->first;
            return p;
        }
        void* p = nullptr;
        hipError_t e = hipMalloc(&p, bytes);
        if (e != hipSuccess) {
            // give cached blocks back and retry once
            trim_locked();
            e = hipMalloc(&p, bytes);
            if (e != hipSuccess) throw HipError(e, "hipMalloc " + std::to_string(bytes) + " bytes");
        }
        size_[p] = bytes;
        live_ += bytes;
        total_ += bytes;
        return p;
    }
    template <class T>
    T* alloc_n(size_t n) {
        return (T*)alloc(sizeof(T) * (n ? n : 1));
    }
    void free(void* p) {
        if (!p) return;
        std::lock_guard<std::mutex> g(mu_);
        auto it = size_.find(p);
        if (it == size_.end()) return;
        free_[it->second].push_back(p);
        live_ -= it->second;
        size_.erase(it);
    }
    void trim() {
        std::lock_guard<std::mutex> g(mu_);
        trim_locked();
    }
    size_t live_bytes() const { return live_; }
    size_t held_bytes() const { return total_; }

  private:
    static size_t round(size_t b) {
        if (b <= (size_t(1) << 20)) {
            size_t q = 256;
            while (q < b) q <<= 1;
            return q;
        }
        return (b + 65535) & ~size_t(65535);  // large sizes: 64 KiB granularity
    }
    void trim_locked() {
        for (auto& kv : free_)
            for (void* p : kv.second) {
                (void)hipFree(p);
                total_ -= kv.first;
            }
        free_.clear();
    }
    void release_all() {
        trim_locked();
        for (auto& kv : size_) (void)hipFree(kv.first);
        size_.clear();
    }
    std::mutex mu_;
    std::map<size_t, std::vector<void*>> free_;
    std::map<void*, size_t> size_;
    size_t live_ = 0, total_ = 0;
};

// RAII device buffer from a DevMem
template <class T>
struct DBuf {
    DevMem* m = nullptr;
    T* p = nullptr;
    size_t n = 0;
    DBuf() = default;
    DBuf(DevMem& mem, size_t count) : m(&mem), p(mem.alloc_n<T>(count)), n(count) {}
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    DBuf(DBuf&& o) noexcept : m(o.m), p(o.p), n(o.n) { o.p = nullptr; }
    DBuf& operator=(DBuf&& o) noexcept {
        reset();
        m = o.m;
        p = o.p;
        n = o.n;
        o.p = nullptr;
        return *this;
    }
    ~DBuf() { reset(); }
    void reset() {
        if (m && p) m->free(p);
        p = nullptr;
        n = 0;
    }
    void alloc(DevMem& mem, size_t count) {
        reset();
        m = &mem;
        p = mem.alloc_n<T>(count);
        n = count;
    }
};

}  // namespace rin
