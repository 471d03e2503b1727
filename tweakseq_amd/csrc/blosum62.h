// BLOSUM62 in the row order A R N D C Q E G H I L K M F P S T W Y V B Z X, as
// tweakseq holds it (tweakseq/Core/Annotations/Consensus.cpp:34-69: the same
// published numbers, by necessity), stored as the lower triangle row by row
// (row r has r + 1 entries).  One text for the host tables (runtime.hip) and
// the device copy the single-pair kernel builds its LDS table from (pair.hip).
#pragma once

#define TA_PROT_ORDER "ARNDCQEGHILKMFPSTWYVBZX"
#define TA_PROT_K 23
#define TA_BLOSUM62_LOWER \
    4, \
    -1, 5, \
    -2, 0, 6, \
    -2, -2, 1, 6, \
    0, -3, -3, -3, 9, \
    -1, 1, 0, 0, -3, 5, \
    -1, 0, 0, 2, -4, 2, 5, \
    0, -2, 0, -1, -3, -2, -2, 6, \
    -2, 0, 1, -1, -3, 0, 0, -2, 8, \
    -1, -3, -3, -3, -1, -3, -3, -4, -3, 4, \
    -1, -2, -3, -4, -1, -2, -3, -4, -3, 2, 4, \
    -1, 2, 0, -1, -3, 1, 1, -2, -1, -3, -2, 5, \
    -1, -1, -2, -3, -1, 0, -2, -3, -2, 1, 2, -1, 5, \
    -2, -3, -3, -3, -2, -3, -3, -3, -1, 0, 0, -3, 0, 6, \
    -1, -2, -2, -1, -3, -1, -1, -2, -2, -3, -3, -1, -2, -4, 7, \
    1, -1, 1, 0, -1, 0, 0, 0, -1, -2, -2, 0, -1, -2, -1, 4, \
    0, -1, 0, -1, -1, -1, -1, -2, -2, -1, -1, -1, -1, -2, -1, 1, 5, \
    -3, -3, -4, -4, -2, -2, -3, -2, -2, -3, -2, -3, -1, 1, -4, -3, -2, 11, \
    -2, -2, -2, -3, -2, -1, -2, -3, 2, -1, -1, -2, -1, 3, -3, -2, -2, 2, 7, \
    0, -3, -3, -3, -1, -2, -2, -3, -3, 3, 1, -2, 1, -1, -2, -2, 0, -3, -1, 4, \
    -2, -1, 3, 4, -3, 0, 1, -1, 0, -3, -4, 0, -3, -3, -2, 0, -1, -4, -3, -3, 4, \
    -1, 0, 0, 1, -3, 3, 4, -2, 0, -3, -3, 1, -1, -3, -1, 0, -1, -3, -2, -2, 1, 4, \
    0, -1, -1, -1, -2, -1, -1, -1, -1, -1, -1, -1, -1, -1, -2, 0, 0, -2, -1, -1, -1, -1, -1
