/*
 * queue.c -- lock-free task counter.
 *
 * Replaces the reference's task queue (blas/blasx_task_queue.c:3-17,
 * include/blasx_task_queue.h:9-13): a shared counter starting at
 * task_num - 1 and handed out downwards by compare-and-swap; -1 when empty.
 * Same order and the same exactly-once guarantee; 64-bit ids.
 */
#include "internal.h"

void bx_queue_init(bx_queue_t *q, int64_t task_num) { q->tail = task_num - 1; }

int64_t bx_queue_next(bx_queue_t *q)
{
    int64_t cur = __atomic_load_n(&q->tail, __ATOMIC_ACQUIRE);
    while (cur >= 0) {
        if (__atomic_compare_exchange_n(&q->tail, &cur, cur - 1, 1, __ATOMIC_ACQ_REL,
                                        __ATOMIC_ACQUIRE))
            return cur;
    }
    return -1;
}
