/* ref_driver.c -- exposes the reference C solver
 * 10Bit_CUDA/004GCC_constellation.c (preset 4) as a library call.
 * TEST INFRASTRUCTURE ONLY (built into oracle/_ref by ref.mk).
 * The reference source is included from where it lies; its main() is
 * renamed so that this file can call genConstellations (:647) and
 * execSolutions (:412) for a single N and read back every constellation
 * record with its weighted count. */
#define main ref004_main
#include REF_SRC
#undef main

/* Returns the record count m; fills up to cap records as
 * {ld, rd, col, startijkl} (u32) and the weighted count per record. */
long ref004_count(int N, unsigned int *rec4, long long *solutions, long cap)
{
    IntHashSet *ijkl = create_int_hashset();
    ConstellationArrayList *cl = create_constellation_arraylist();
    genConstellations(ijkl, cl, N);
    execSolutions(cl, N);
    long m = cl->size;
    for (long a = 0; a < m && a < cap; a++) {
        Constellation *c = &cl->data[a];
        rec4[4 * a + 0] = (unsigned int)c->ld;
        rec4[4 * a + 1] = (unsigned int)c->rd;
        rec4[4 * a + 2] = (unsigned int)c->col;
        rec4[4 * a + 3] = (unsigned int)c->startijkl;
        solutions[a] = c->solutions;
    }
    free_int_hashset(ijkl);
    free_constellation_arraylist(cl);
    return m;
}
