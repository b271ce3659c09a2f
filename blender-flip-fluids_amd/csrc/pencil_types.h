// pencil_types.h -- plain types shared by the line-pipeline sweep
// (pencil_sweep.h), its host launchers (kernels.h) and its CPU emulation.
#pragma once

namespace ff {

// Rows of one 32-cell word of a (j, k) line along i: cell i of the word is a
// pressure row when bit i&31 is set, and its row is base + the number of set
// bits below it (base = the row of the word's first liquid cell in k,j,i
// order, i.e. the exclusive scan of the row mask at the word's first cell)
struct LineWord {
    int base;
    unsigned bits;
};

}  // namespace ff
