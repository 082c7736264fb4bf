/*
 * Ring of page-locked wave-function frames handed from the thread that runs
 * the engine to a writer thread.  Same names and use as the reference's
 * dev/gpe_queue.h:46-139 (create_gpe_queue, get_gpe_queue_buffer,
 * add_to_gpe_queue, take_off_gpe_queue, destroy_gpe_queue, the global
 * wf_queue with member q, and flag_stop), with two differences a driver does
 * not see: q is a locked queue, safe to use from both threads (the reference
 * shares a std::queue without a lock), and flag_stop is atomic.
 * add_to_gpe_queue blocks while the ring is full instead of spinning.
 * Frames are allocated with gpe_host_alloc (libgpe_hip.so).  C++ only, as in
 * the reference.
 */
#ifndef GPE_QUEUE_H
#define GPE_QUEUE_H

#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include <atomic>
#include <condition_variable>
#include <mutex>
#include <queue>

#include "gpe_engine.h"

/* tells the writer thread that no more frames will come */
inline std::atomic<uint8_t> flag_stop{0};

/* std::queue interface behind one mutex */
template <typename T>
class gpe_locked_queue {
  public:
    bool empty() const {
        std::lock_guard<std::mutex> lk(mu_);
        return q_.empty();
    }
    size_t size() const {
        std::lock_guard<std::mutex> lk(mu_);
        return q_.size();
    }
    void push(const T& v) {
        {
            std::lock_guard<std::mutex> lk(mu_);
            q_.push(v);
        }
        cv_.notify_all();
    }
    T front() const {
        std::lock_guard<std::mutex> lk(mu_);
        return q_.front();
    }
    void pop() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            q_.pop();
        }
        cv_.notify_all();
    }
    /* pushes once fewer than limit elements are queued */
    void push_bounded(const T& v, size_t limit) {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return q_.size() < limit; });
        q_.push(v);
        lk.unlock();
        cv_.notify_all();
    }
    /* removes the front element; false if empty */
    bool try_take(T* out) {
        std::lock_guard<std::mutex> lk(mu_);
        if (q_.empty()) return false;
        *out = q_.front();
        q_.pop();
        cv_.notify_all();
        return true;
    }

  private:
    mutable std::mutex mu_;
    std::condition_variable cv_;
    std::queue<T> q_;
};

typedef struct {
    gpe_locked_queue<cplx *> q;
    uint16_t counter;
    uint16_t max_size;
    cplx **data;
} gpe_queue;

/* defined by the driver, as in the reference */
extern gpe_queue wf_queue;

/* allocates queue_size frames of nxyz complex128 values (at most 5); returns
 * the single frame when queue_size is 1, NULL otherwise */
static inline void *create_gpe_queue(const uint64_t nxyz, uint16_t queue_size = 5)
{
    const uint16_t queue_max_size = 5;
    if (queue_size > queue_max_size) {
        printf("Error: Queue buffer is too big! Truncating queue size to %d\n", queue_max_size);
        queue_size = queue_max_size;
    }
    if (queue_size < 1) queue_size = 1;
    printf("GPE Queue size (wavefunctions to be buffered at once): %u\n", queue_size);
    wf_queue.counter = 0;
    wf_queue.max_size = queue_size;
    flag_stop = 0;
    wf_queue.data = (cplx **)malloc(queue_size * sizeof(cplx *));
    for (uint16_t i = 0; i < queue_size; i++) {
        if (gpe_host_alloc((void **)&wf_queue.data[i], sizeof(cplx) * nxyz) != 0) {
            printf("Error: Cannot allocate memory for gpe_queue!\n");
            exit(EXIT_FAILURE);
        }
    }
    return (queue_size == 1) ? (void *)wf_queue.data[0] : NULL;
}

static inline cplx *get_gpe_queue_buffer(void) { return wf_queue.data[wf_queue.counter % wf_queue.max_size]; }

/* queues a filled frame; waits while max_size - 1 frames are pending, so the
 * frame handed out next by get_gpe_queue_buffer is not one still queued */
static inline void add_to_gpe_queue(cplx *psi)
{
    wf_queue.q.push_bounded(psi, (size_t)(wf_queue.max_size > 1 ? wf_queue.max_size - 1 : 1));
    wf_queue.counter++;
}

static inline cplx *take_off_gpe_queue(void)
{
    cplx *psi = NULL;
    wf_queue.q.try_take(&psi);
    return psi;
}

static inline void destroy_gpe_queue(void)
{
    for (uint16_t i = 0; i < wf_queue.max_size; i++) gpe_host_free(wf_queue.data[i]);
    flag_stop = 1;
    free(wf_queue.data);
    wf_queue.data = NULL;
}

#endif
