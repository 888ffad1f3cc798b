// Forced-include configuration for building the LoAdSG reference sources at a
// chosen dimension.  The reference fixes its dimension with a macro in its own
// configuration header (library/source/abbrevi.h:11-15, guarded by ABBREVI_H).
// Defining the guard and the same four macros here, before any reference file
// is read (g++ -include), selects the dimension without touching or copying any
// reference file.  LSG_REF_DIM is passed on the compiler command line.
#ifndef ABBREVI_H
#define ABBREVI_H
#ifndef LSG_REF_DIM
#error "LSG_REF_DIM must be defined"
#endif
#define DimensionSparseGrid LSG_REF_DIM
#define MaximumDepth 15
#define indexInteger int
#define dataInteger unsigned long
#define testNow
#endif
