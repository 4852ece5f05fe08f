/*
 * msgs_sort.h -- restatement of libstdc++'s std::sort (GCC 11, bits/stl_algo.h and
 * bits/stl_heap.h: introsort with median-of-three pivot, unguarded partition, heapsort
 * fallback at depth 2*floor(log2 n), final insertion sort with threshold 16).
 *
 * The reference sorts SMEMs by loglik (src/SMEM.cpp:275-276), seeds by operator<
 * (src/SeedPair.cpp:75) and alignments by log10lik (src/Alignment.h:629-633) with
 * std::sort, which is not stable: which of two equal keys comes first is decided by
 * the exact sequence of compares and moves.  This header repeats that sequence so a
 * sort on the device (or anywhere else) yields the same permutation as the
 * reference's host sort for any input.  The recursion of __introsort_loop is replaced
 * by an explicit stack; the sub-ranges it visits are disjoint, so the order in which
 * they are processed does not change the result.
 *
 * Usable from host and device code (HD).
 */
#ifndef MSGS_SORT_H
#define MSGS_SORT_H

#include <stdint.h>

#if defined(__HIPCC__)
#define MSGS_HD __host__ __device__
#else
#define MSGS_HD
#endif

namespace msgs_sort {

template <class T> MSGS_HD inline void swapv(T& a, T& b) { T t = a; a = b; b = t; }

MSGS_HD inline long lg(long n) { return 63 - __builtin_clzll((unsigned long long)n); }

template <class T, class Cmp> MSGS_HD void move_median_to_first(T* result, T* a, T* b, T* c, Cmp& comp) {
    if (comp(*a, *b)) {
        if (comp(*b, *c)) swapv(*result, *b);
        else if (comp(*a, *c)) swapv(*result, *c);
        else swapv(*result, *a);
    } else if (comp(*a, *c)) swapv(*result, *a);
    else if (comp(*b, *c)) swapv(*result, *c);
    else swapv(*result, *b);
}

template <class T, class Cmp> MSGS_HD T* unguarded_partition(T* first, T* last, T* pivot, Cmp& comp) {
    while (true) {
        while (comp(*first, *pivot)) ++first;
        --last;
        while (comp(*pivot, *last)) --last;
        if (!(first < last)) return first;
        swapv(*first, *last);
        ++first;
    }
}

template <class T, class Cmp> MSGS_HD T* unguarded_partition_pivot(T* first, T* last, Cmp& comp) {
    T* mid = first + (last - first) / 2;
    move_median_to_first(first, first + 1, mid, last - 1, comp);
    return unguarded_partition(first + 1, last, first, comp);
}

template <class T, class Cmp> MSGS_HD void push_heap(T* first, long hole, long top, T value, Cmp& comp) {
    long parent = (hole - 1) / 2;
    while (hole > top && comp(first[parent], value)) {
        first[hole] = first[parent];
        hole = parent;
        parent = (hole - 1) / 2;
    }
    first[hole] = value;
}

template <class T, class Cmp> MSGS_HD void adjust_heap(T* first, long hole, long len, T value, Cmp& comp) {
    const long top = hole;
    long second = hole;
    while (second < (len - 1) / 2) {
        second = 2 * (second + 1);
        if (comp(first[second], first[second - 1])) second--;
        first[hole] = first[second];
        hole = second;
    }
    if ((len & 1) == 0 && second == (len - 2) / 2) {
        second = 2 * (second + 1);
        first[hole] = first[second - 1];
        hole = second - 1;
    }
    push_heap(first, hole, top, value, comp);
}

template <class T, class Cmp> MSGS_HD void make_heap(T* first, T* last, Cmp& comp) {
    const long len = last - first;
    if (len < 2) return;
    long parent = (len - 2) / 2;
    while (true) {
        T value = first[parent];
        adjust_heap(first, parent, len, value, comp);
        if (parent == 0) return;
        parent--;
    }
}

template <class T, class Cmp> MSGS_HD void pop_heap(T* first, T* last, T* result, Cmp& comp) {
    T value = *result;
    *result = *first;
    adjust_heap(first, 0, (long)(last - first), value, comp);
}

template <class T, class Cmp> MSGS_HD void partial_sort_all(T* first, T* last, Cmp& comp) {
    /* __partial_sort(first, last, last): __heap_select (make_heap; no elements past
     * middle) then __sort_heap */
    make_heap(first, last, comp);
    while (last - first > 1) {
        --last;
        pop_heap(first, last, last, comp);
    }
}

template <class T, class Cmp> MSGS_HD void unguarded_linear_insert(T* last, Cmp& comp) {
    T val = *last;
    T* next = last - 1;
    while (comp(val, *next)) {
        *last = *next;
        last = next;
        --next;
    }
    *last = val;
}

template <class T, class Cmp> MSGS_HD void insertion_sort(T* first, T* last, Cmp& comp) {
    if (first == last) return;
    for (T* i = first + 1; i != last; ++i) {
        if (comp(*i, *first)) {
            T val = *i;
            for (T* k = i; k != first; --k) *k = *(k - 1);
            *first = val;
        } else {
            unguarded_linear_insert(i, comp);
        }
    }
}

template <class T, class Cmp> MSGS_HD void sort(T* first, T* last, Cmp comp) {
    if (first == last) return;
    struct Frame { T* f; T* l; long d; };
    Frame stack[72];
    int sp = 0;
    stack[sp++] = Frame{first, last, 2 * lg((long)(last - first))};
    while (sp > 0) {
        Frame fr = stack[--sp];
        T* f = fr.f;
        T* l = fr.l;
        long d = fr.d;
        while (l - f > 16) {
            if (d == 0) { partial_sort_all(f, l, comp); break; }
            --d;
            T* cut = unguarded_partition_pivot(f, l, comp);
            stack[sp++] = Frame{cut, l, d};
            l = cut;
        }
    }
    /* __final_insertion_sort */
    if (last - first > 16) {
        insertion_sort(first, first + 16, comp);
        for (T* i = first + 16; i != last; ++i) unguarded_linear_insert(i, comp);
    } else {
        insertion_sort(first, last, comp);
    }
}

}  // namespace msgs_sort
#endif
