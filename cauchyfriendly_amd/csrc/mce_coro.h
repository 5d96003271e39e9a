// A minimal lazily started coroutine task for the host side of an update.
//
// An update is written as a chain of Co<T> coroutines that await each other.  Run on its
// own (run_sync), nothing in it ever suspends: every device read completes where it is
// awaited, so the chain runs straight through like ordinary calls.  Run by a batch
// scheduler, a device read suspends the chain (the innermost task records itself as the
// one to resume) and control returns to the scheduler, which resumes it once the read's
// data is on the host.  A finished task resumes the task that awaited it.
#pragma once
#include <coroutine>
#include <exception>
#include <stdexcept>
#include <utility>

namespace mce {

struct CoPromiseBase {
    std::coroutine_handle<> cont;  // the awaiting task, if any
    std::exception_ptr exc;
    std::suspend_always initial_suspend() noexcept { return {}; }
    struct Final {
        bool await_ready() noexcept { return false; }
        template <class P>
        std::coroutine_handle<> await_suspend(std::coroutine_handle<P> h) noexcept {
            const std::coroutine_handle<> c = h.promise().cont;
            return c ? c : std::noop_coroutine();
        }
        void await_resume() noexcept {}
    };
    Final final_suspend() noexcept { return {}; }
    void unhandled_exception() { exc = std::current_exception(); }
};

template <class T>
struct CoPromise : CoPromiseBase {
    T val{};
    void return_value(T v) { val = std::move(v); }
    T take() { return std::move(val); }
};
template <>
struct CoPromise<void> : CoPromiseBase {
    void return_void() {}
    void take() {}
};

template <class T = void>
class [[nodiscard]] Co {
   public:
    struct promise_type : CoPromise<T> {
        Co get_return_object() { return Co(std::coroutine_handle<promise_type>::from_promise(*this)); }
    };
    Co(Co&& o) noexcept : h_(std::exchange(o.h_, {})) {}
    Co& operator=(Co&& o) noexcept {
        if (this != &o) {
            if (h_) h_.destroy();
            h_ = std::exchange(o.h_, {});
        }
        return *this;
    }
    Co(const Co&) = delete;
    Co& operator=(const Co&) = delete;
    ~Co() {
        if (h_) h_.destroy();
    }

    // awaited by another task: start this one, resume the awaiting one when it ends
    bool await_ready() const noexcept { return false; }
    std::coroutine_handle<> await_suspend(std::coroutine_handle<> awaiting) noexcept {
        h_.promise().cont = awaiting;
        return h_;
    }
    T await_resume() { return result(); }

    // top-level use by a scheduler
    void start() { h_.resume(); }
    bool done() const { return h_.done(); }
    T result() {
        if (h_.promise().exc) std::rethrow_exception(h_.promise().exc);
        return h_.promise().take();
    }

   private:
    explicit Co(std::coroutine_handle<promise_type> h) : h_(h) {}
    std::coroutine_handle<promise_type> h_;
};

// Runs a task that never suspends to its end and returns its result.
template <class T>
T run_sync(Co<T>&& c) {
    c.start();
    if (!c.done()) throw std::logic_error("a task suspended outside a batch scheduler");
    return c.result();
}

}  // namespace mce
