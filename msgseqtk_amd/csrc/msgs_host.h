/*
 * msgs_host.h -- host-side state of libmsgs_hip.so (context, resident database,
 * resident read batches) shared by its translation units.
 */
#ifndef MSGS_HOST_H
#define MSGS_HOST_H

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/msgs_hip.h"
#include "msgs_kernels.h"

struct DbState {
    int64_t n = 0;
    int64_t B[256] = {}, C[256] = {};
    double ll[16] = {};
    std::vector<std::string> names;
    std::vector<int64_t> chromLen, chromBD, tgtOff;
    int32_t nChrom = 0;
    /* device */
    uint8_t* dBd = nullptr;
    /* narrow index (n < 2^31): uint32_t rows and absolute occ-block counts; wide index
     * (n < 2^34, or every index with the context option wide_index): int64_t rows and
     * occ-block counts from the start of a superblock of 2^supShift positions */
    bool wide = false;
    int32_t supShift = 31;
    void* dSa = nullptr;
    void* dIsa = nullptr;
    int64_t* dSup = nullptr;   /* wide: A, C, G, T counts before each superblock */
    /* the reference's block-by-block build (msgseqtk-build with a block size below the BD
     * length): BD end of each block, in text order; empty for one block */
    std::vector<int64_t> blockEnds;
    uint8_t* dBwt = nullptr;
    OccBlock* dOcc = nullptr;
    int64_t* dNPos = nullptr;
    int64_t nN = 0;
    int64_t* dChromBD = nullptr;
    int64_t* dTgtOff = nullptr;
    int64_t* dChromLen = nullptr;
    uint8_t* dTgt = nullptr;
    std::vector<uint8_t> tgtHost;   /* chrom sequences + terminators (for record text) */
    void release();
    DevIndex view() const;
};

/* suffix array, ISA, BWT and occ blocks of db.dBd; wide: build the wide layout (set by the
 * caller; required from 2^31 symbols on) with superblocks of 2^supShift positions and
 * suffix-sorting ranges of at most rangeMax suffixes (the test shape uses small ones) */
int buildIndexDevice(DbState& db, hipStream_t st, std::string& err, int64_t rangeMax);

#endif
