// scmf_batch_tickets.h — steps of a batch behind a ticket (library-internal,
// for the job server): events around the enqueued steps on the batch stream,
// read without waiting, with their end times on the host clock.
#pragma once
#include <stdint.h>

#include <chrono>
#include <mutex>

#include "../../include/scmf.h"

namespace scmf_batch_tickets {

// Enqueue n_steps steps of the batch (as scmf_batch_step) between two events.
int enqueue(scmf_batch* b, int32_t n_steps, uint64_t* ticket);
// *done = 1 once the steps of `ticket` have run (wait: block until they have);
// then *t_end = when they ended, in seconds after origin(b), and *busy = their
// device time; the ticket is forgotten.
int poll(scmf_batch* b, uint64_t ticket, bool wait, int* done, double* t_end, double* busy);
// Host time of the batch's origin event.
std::chrono::steady_clock::time_point origin(scmf_batch* b);
// Held around every graph capture of a batch: a thread that creates or
// destroys systems while another may be capturing (the job server's worker)
// takes it too, since the runtime abandons a capture when another thread
// synchronises or frees meanwhile.
std::mutex& capture_mutex();

}  // namespace scmf_batch_tickets
