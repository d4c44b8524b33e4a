/*
 * spike_pstrsv for AMD Instinct MI355X -- global variables read by the
 * reference's timing harness.
 *
 * Replaces the harness-facing part of the reference's src/spike_pstrsv_globals.h
 * (declarations at :53-79; definitions in src/spike_pstrsv.c:18-37 and
 * src/spike_pstrsv_ops.c:18-67), so that a driver written against the
 * reference (test/runtime_analysis.c includes this header and reads these
 * names) compiles and links unchanged.  The reference's matrix-sized static
 * arrays (D, T, T2, ...) are internal to its CPU implementation and are not
 * provided.
 *
 * Meaning of the timers here (seconds, accumulated like the reference's;
 * reset by _spike_pstrsv_preproc, filled by the profiling twins
 * _spike_pstrsv_preproc / _spike_pstrsv_solve only):
 *   preprocessD_timer_pstrsv3           host symbolic analysis (validation,
 *                                       partitions, dependents, block-solve
 *                                       schedules; the reference's preprocessD,
 *                                       setDependents, preprocessS, preprocessR
 *                                       and preprocessMappings steps)
 *   produceS_timer_pstrsv3              spike build on the device
 *   produceReducedSystem_timer_pstrsv3  reduced-system assembly and its solver
 *                                       set-up
 *   reduced_solver_timer_pstrsv3        reduced solve (device time)
 *   computeY_timer[0]                   pass-1 block solve (device time)
 *   subtract_Rx_timer[0]                coupling update (device time)
 *   solveD2_timer[0]                    pass-2 block solve (device time)
 * All partitions of one device phase run in one launch, so per-thread slots
 * other than [0] stay 0; the remaining timers (memoryAlloc, produceCompressedR,
 * waitBefore/waitAfter, solveIndependent, solveD, init_vars) have no separate
 * phase here and stay 0.
 *
 * PSTRSV_VERBOSE != 0 makes spike_pstrsv_preproc print the partition log in
 * the reference's logPSTRSV3 layout.  FORCE_LOAD_BALANCE and
 * IS_MATRIX_PARTITIONED are accepted and ignored (caller partition boundaries
 * are available through spike_factor_create_partitioned).
 */
#ifndef SPIKE_PSTRSV_GLOBALS_H
#define SPIKE_PSTRSV_GLOBALS_H

#define MAXTHREADCOUNT 128

#ifdef __cplusplus
extern "C" {
#endif

extern char PSTRSV_VERBOSE;
extern char FORCE_LOAD_BALANCE;
extern char IS_MATRIX_PARTITIONED;

extern double memoryAlloc_timer_pstrsv3;
extern double preprocessD_timer_pstrsv3;
extern double preprocessR_timer_pstrsv3;
extern double setDependents_timer_pstrsv3;
extern double preprocessS_timer_pstrsv3;
extern double preprocessMappings_timer_pstrsv3;
extern double produceCompressedR_timer_pstrsv3;
extern double produceS_timer_pstrsv3;
extern double produceReducedSystem_timer_pstrsv3;
extern double init_vars_pstrsv3;
extern double reduced_solver_timer_pstrsv3;
extern double temp_timer[MAXTHREADCOUNT];
extern double waitBefore_timer_pstrsv3[MAXTHREADCOUNT];
extern double computeY_timer[MAXTHREADCOUNT];
extern double solveIndependent_timer[MAXTHREADCOUNT];
extern double subtract_Rx_timer[MAXTHREADCOUNT];
extern double waitAfter_timer_pstrsv3[MAXTHREADCOUNT];
extern double solveD2_timer[MAXTHREADCOUNT];
extern double solveD_timer[MAXTHREADCOUNT];

#ifdef __cplusplus
}
#endif

#endif /* SPIKE_PSTRSV_GLOBALS_H */
