/*
 * Ring of page-locked wave-function frames handed from the thread that runs
 * the engine to a writer thread.  Same names and use as the reference's
 * dev/gpe_queue.h:46-139 (create_gpe_queue, get_gpe_queue_buffer,
 * add_to_gpe_queue, take_off_gpe_queue, destroy_gpe_queue, the global
 * wf_queue with member q, and flag_stop), with differences a driver does not
 * see: q is a locked queue, safe to use from both threads (the reference
 * shares a std::queue without a lock), flag_stop is atomic, and
 * get_gpe_queue_buffer hands out only a frame that is neither queued nor
 * still being written out (the reference's ring index can return the frame
 * the writer thread is writing), waiting for the writer instead of spinning.
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
#include <vector>

#include "gpe_engine.h"

/* tells the writer thread that no more frames will come */
inline std::atomic<uint8_t> flag_stop{0};

/* Frames come from gpe_host_alloc (page-locked, libgpe_hip.so) unless the
 * including file defines GPE_QUEUE_HOST_ALLOC / GPE_QUEUE_HOST_FREE. */
#ifndef GPE_QUEUE_HOST_ALLOC
#define GPE_QUEUE_HOST_ALLOC(pp, bytes) gpe_host_alloc((pp), (bytes))
#define GPE_QUEUE_HOST_FREE(p) gpe_host_free(p)
#endif

/* The frame queue: the std::queue interface the drivers use (empty, size,
 * front, pop, push) behind one mutex, plus the state of every frame of the
 * ring, so that a frame is handed to the producer only when it is neither
 * queued nor still being written by the consumer.  The consumer holds the
 * frame it took last until it calls into the queue again (empty, front, pop,
 * try_take: the reference's writer loop does so after each fwrite) or
 * releases it with release_gpe_queue_frame(). */
template <typename T>
class gpe_locked_queue {
  public:
    enum : uint8_t { FREE = 0, FILLING = 1, QUEUED = 2, HELD = 3 };

    void set_frames(T *frames, size_t n) {
        std::lock_guard<std::mutex> lk(mu_);
        frames_.assign(frames, frames + n);
        state_.assign(n, FREE);
        std::queue<T>().swap(q_);
    }
    bool empty() {
        std::lock_guard<std::mutex> lk(mu_);
        release_held_locked();
        return q_.empty();
    }
    size_t size() const {
        std::lock_guard<std::mutex> lk(mu_);
        return q_.size();
    }
    void push(const T &v) {
        {
            std::lock_guard<std::mutex> lk(mu_);
            set_state_locked(v, QUEUED);
            q_.push(v);
        }
        cv_.notify_all();
    }
    T front() {
        std::lock_guard<std::mutex> lk(mu_);
        release_held_locked();
        return q_.front();
    }
    void pop() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            release_held_locked();
            set_state_locked(q_.front(), HELD);
            q_.pop();
        }
        cv_.notify_all();
    }
    /* removes the front element (held by the caller until its next call); false if empty */
    bool try_take(T *out) {
        std::lock_guard<std::mutex> lk(mu_);
        release_held_locked();
        if (q_.empty()) return false;
        *out = q_.front();
        set_state_locked(*out, HELD);
        q_.pop();
        cv_.notify_all();
        return true;
    }
    /* a frame that is neither queued nor held, marked as being filled;
     * waits while there is none */
    T acquire() {
        std::unique_lock<std::mutex> lk(mu_);
        for (;;) {
            for (size_t i = 0; i < frames_.size(); i++) {
                const size_t k = (next_ + i) % frames_.size();
                if (state_[k] == FREE) {
                    state_[k] = FILLING;
                    next_ = k + 1;
                    return frames_[k];
                }
            }
            cv_.wait(lk);
        }
    }
    /* the consumer is done with the frame it took last */
    void release_held() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            release_held_locked();
        }
        cv_.notify_all();
    }

  private:
    void set_state_locked(const T &v, uint8_t st) {
        for (size_t i = 0; i < frames_.size(); i++)
            if (frames_[i] == v) state_[i] = st;
    }
    void release_held_locked() {
        bool any = false;
        for (auto &st : state_)
            if (st == HELD) {
                st = FREE;
                any = true;
            }
        if (any) cv_.notify_all();
    }
    mutable std::mutex mu_;
    std::condition_variable cv_;
    std::queue<T> q_;
    std::vector<T> frames_;
    std::vector<uint8_t> state_;
    size_t next_ = 0;
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
        if (GPE_QUEUE_HOST_ALLOC((void **)&wf_queue.data[i], sizeof(cplx) * nxyz) != 0) {
            printf("Error: Cannot allocate memory for gpe_queue!\n");
            exit(EXIT_FAILURE);
        }
    }
    wf_queue.q.set_frames(wf_queue.data, queue_size);
    return (queue_size == 1) ? (void *)wf_queue.data[0] : NULL;
}

/* a frame to fill: one that is neither queued nor still held by the writer
 * (waits for the writer when every frame is in use) */
static inline cplx *get_gpe_queue_buffer(void) { return wf_queue.q.acquire(); }

/* queues a filled frame */
static inline void add_to_gpe_queue(cplx *psi)
{
    wf_queue.q.push(psi);
    wf_queue.counter++;
}

/* the oldest queued frame, or NULL; it stays the writer's until the writer
 * calls into the queue again */
static inline cplx *take_off_gpe_queue(void)
{
    cplx *psi = NULL;
    wf_queue.q.try_take(&psi);
    return psi;
}

/* the writer is done with the frame it took last (optional: its next call
 * into the queue does the same) */
static inline void release_gpe_queue_frame(void) { wf_queue.q.release_held(); }

static inline void destroy_gpe_queue(void)
{
    for (uint16_t i = 0; i < wf_queue.max_size; i++) GPE_QUEUE_HOST_FREE(wf_queue.data[i]);
    flag_stop = 1;
    free(wf_queue.data);
    wf_queue.data = NULL;
}

#endif
