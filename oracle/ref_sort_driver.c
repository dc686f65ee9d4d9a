/*
 * ref_sort_driver.c -- TEST INFRASTRUCTURE ONLY.
 *
 * Thin C entry point around the reference's own local presort
 * (sl_pepc src/core/sort_radix.c:129-151 sort_radix_ip, with sort_insert.c),
 * compiled from the sources under /root/reference by oracle/Makefile into
 * oracle/_ref/libref_sort.so.  tests/test_oracle_sort.py uses it to pin the
 * oracle's restatement, including the order it gives to equal keys.
 * Parameters as slsort_keys passes them (mpi_pepckeys.c:29-31, :167):
 * rhigh = rlow = rwidth = -1, i.e. bits 63..0 in 10-bit digits.
 */
#include "sl_common.h"

int ref_sort_radix_ip(long long n, long long *keys, long long *idx, double *work)
{
    elements_t s;
    memset(&s, 0, sizeof(s));
    elem_set_size(&s, n);
    elem_set_max_size(&s, n);
    elem_set_keys(&s, (slkey_t *)keys);
    elem_set_indices(&s, (slindex_t *)idx);
    s.data0 = work;
    return (int)sort_radix_ip(&s, NULL, -1, -1, -1);
}

int ref_sizeof_key(void) { return (int)sizeof(slkey_t); }
int ref_sizeof_index(void) { return (int)sizeof(slindex_t); }
