// Host atomics of the drop-in front end (CAS, writeMin, writeAdd; the
// reference's utils.h:423-477) hammered from several OpenMP threads on
// 4- and 8-byte element types.  Each array carries guard elements on both
// sides of the target, which must stay untouched.  Prints one line per type
// and exits non-zero on the first wrong sum, minimum or guard.
#define GG_FRONT_NO_DRIVER
#include "ligra-numa.h"
#undef main

#include <omp.h>

#include <cstdio>

template <class T>
static int check_type(const char *name, int threads, int reps)
{
    const T guard = T(-7);
    // target in the middle, then at the end of the array (the last element
    // is where a wider CAS would run past the allocation)
    for (int pos : {1, 3}) {
        T *a = new T[4];
        for (int i = 0; i < 4; i++) a[i] = guard;
        a[pos] = T(0);
#pragma omp parallel num_threads(threads)
        for (int r = 0; r < reps; r++) writeAdd(&a[pos], T(1));
        const T want = T(threads) * T(reps);
        if (a[pos] != want) {
            std::printf("%s writeAdd pos %d: %g != %g\n", name, pos, (double)a[pos], (double)want);
            return 1;
        }
        for (int i = 0; i < 4; i++)
            if (i != pos && a[i] != guard) {
                std::printf("%s writeAdd guard %d overwritten: %g\n", name, i, (double)a[i]);
                return 1;
            }
        // writeMin: every thread offers its own values; the minimum wins
        a[pos] = T(1000000);
        int wins = 0;
#pragma omp parallel num_threads(threads) reduction(+ : wins)
        {
            const int t = omp_get_thread_num();
            for (int k = 0; k < reps / 16; k++) wins += writeMin(&a[pos], T(3 + t + (k * 7919) % 5000));
        }
        if (a[pos] != T(3) || wins < 1) {
            std::printf("%s writeMin pos %d: %g (wins %d)\n", name, pos, (double)a[pos], wins);
            return 1;
        }
        // CAS: exactly one thread moves each value on by one
        a[pos] = T(0);
        int moved = 0;
#pragma omp parallel num_threads(threads) reduction(+ : moved)
        for (int r = 0; r < reps; r++) {
            const T seen = atomic_peek(&a[pos]);
            moved += CAS(&a[pos], seen, T(seen + T(1)));
        }
        if (a[pos] != T(moved)) {
            std::printf("%s CAS pos %d: %g after %d successes\n", name, pos, (double)a[pos], moved);
            return 1;
        }
        for (int i = 0; i < 4; i++)
            if (i != pos && a[i] != guard) {
                std::printf("%s guard %d overwritten: %g\n", name, i, (double)a[i]);
                return 1;
            }
        delete[] a;
    }
    std::printf("%s ok\n", name);
    return 0;
}

int main()
{
    const int threads = 8, reps = 20000;
    int rc = 0;
    rc |= check_type<float>("float", threads, reps);
    rc |= check_type<int>("int", threads, reps);
    rc |= check_type<unsigned>("unsigned", threads, reps);
    rc |= check_type<double>("double", threads, reps);
    rc |= check_type<long>("long", threads, reps);
    return rc;
}
