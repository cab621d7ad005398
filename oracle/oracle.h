/*
 * oracle.h -- CPU oracle for the SPID-Join equi-join path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this code; only tests/, __graft_entry__.smoke() and the cpu_baseline
 * leg of bench.py use it, and only as the checker / reported CPU baseline.
 *
 * Contents
 *   1. The reference's hash functions, restated from their published
 *      algorithms and the reference's constants (src/dpu/src/hash.c).
 *   2. The closed-form join oracle: the exact multiset of (R rid, S rid)
 *      pairs of a PK-FK equi-join, plus an order-independent checksum.
 *   3. A multithreaded CPU restatement of the reference join algorithm
 *      (global SuperFastHash partition into D = 64*ranks virtual DPUs,
 *      per-DPU local partition with P_loc = 2^bitlen(n/1935), 4096-slot
 *      linear-probe build / first-match probe, the SPID bank-set / rank-set
 *      replication mapping, the sort-merge and nested-loop variants).
 *
 * Parity pinning (what checks each function; see tests/test_oracle.py):
 *   or_crc_table           -- CRC-32 polynomial closed form; spot values;
 *                             byte-equal to the table in the reference build
 *                             (oracle/_ref) when that is built.
 *   or_glb_partition_hash, or_join_hash, or_local_partition_hash,
 *   or_join_hash2          -- known-answer vectors in tests/golden/ written
 *                             from the reference's own hash.c compiled by
 *                             oracle/Makefile into oracle/_ref.
 *   or_join_expected       -- brute force on tiny inputs; the inverse-
 *                             permutation closed form.
 *   or_spid_join_cpu       -- sorted multiset equals or_join_expected on
 *                             seeded inputs, all algorithms / designs.
 */
#ifndef SPJ_ORACLE_H
#define SPJ_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { uint32_t key; uint32_t rid; } or_tuple_t;   /* kv_pair_t: type_defs.hpp:53-56 */
typedef struct { uint32_t r_rid; uint32_t s_rid; } or_pair_t; /* a2a_srj_probe_inner_mt.c:142-143 */

/* 1. hashes ---------------------------------------------------------------- */
const uint32_t *or_crc_table(void);
uint32_t or_glb_partition_hash(uint32_t key);
uint32_t or_join_hash(uint32_t key);
uint32_t or_local_partition_hash(uint32_t key);
uint32_t or_join_hash2(uint32_t key);
void or_hash_batch(int which, const uint32_t *keys, uint32_t *out, uint64_t n);

/* 2. closed-form oracle ----------------------------------------------------- */
/* Writes the sorted (by r_rid, then s_rid) pairs; returns the pair count, or
 * -1 when out_cap is too small, -2 on allocation failure.  Keys equal to 0 are
 * "empty" (never joined), as in the reference. */
int64_t or_join_expected(const or_tuple_t *R, uint64_t nR, const or_tuple_t *S, uint64_t nS,
                         or_pair_t *out, uint64_t out_cap);
/* Order-independent checksum of a pair multiset: sum and xor of a 64-bit
 * mix of each (r_rid, s_rid). */
void or_pairs_checksum(const or_pair_t *pairs, uint64_t n, uint64_t *sum, uint64_t *xr);
/* Checksum of the expected output without materialising it. */
int or_join_checksum(const or_tuple_t *R, uint64_t nR, const or_tuple_t *S, uint64_t nS,
                     int threads, uint64_t *count, uint64_t *sum, uint64_t *xr);
/* Sort pairs in place by (r_rid, s_rid). */
void or_sort_pairs(or_pair_t *pairs, uint64_t n);

/* 3. CPU restatement of the reference join ---------------------------------- */
typedef struct {
    int rank_num;   /* ranks; D = 64*rank_num virtual DPUs     (idpHandler.hpp:134) */
    int rank_set;   /* G rank groups                             (idpHandler.hpp:136) */
    int bank_set;   /* B DPUs per bank set; 1 = plain path       (idpHandler.hpp:137) */
    int join_alg;   /* 0 phj, 2 smj, 3 nl                        (idpHandler.hpp:126-129) */
    int threads;    /* worker threads (0 = all online cores) */
} or_design_t;

/* Returns the number of pairs written (unsorted, per-DPU order), or a
 * negative error: -1 out_cap too small, -2 allocation, -3 a local partition
 * exceeded the 4096-slot table, -4 duplicate R key, -5 bad design. */
int64_t or_spid_join_cpu(const or_tuple_t *R, uint64_t nR, const or_tuple_t *S, uint64_t nS,
                         const or_design_t *design, or_pair_t *out, uint64_t out_cap);

#ifdef __cplusplus
}
#endif
#endif
