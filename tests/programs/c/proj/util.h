/* Project header of a two-file program: a prototype, a macro and a global declared here
 * and defined in util.c. */
#ifndef UTIL_H
#define UTIL_H
#define RING_TAG 7
extern int hops;
int ring_pass(int value, int rank, int size);
#endif
