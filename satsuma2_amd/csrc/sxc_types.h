// sxc_types.h -- plain data shared by the host library, the kernels and the CPU
// workgroup emulator.  No HIP or C++ runtime types: every struct here crosses the
// host/device boundary by memcpy.
#pragma once
#include <stdint.h>

#define SXC_MIN_NB 9   // N = 512
#define SXC_MAX_NB 15  // N = 32768
#define SXC_ENV_SIZE 256        // SeqAnalyzer::m_envSize (analysis/CrossCorr.cc:562)
#define SXC_MIN_LEN 45          // SeqAnalyzer::m_minLen (analysis/CrossCorr.cc:560)
#define SXC_WIN_THRESHOLD 1889  // (int)(45 * 0.42 * 100) (analysis/CrossCorr.cc:675)
#define SXC_NCLASS 18           // 16 IUPAC letters + '-' + "zero row" class
#define SXC_CLASS_N 14          // class id of 'N'
#define SXC_CLASS_ZERO 17       // bytes with an all-zero codec row
#define SXC_PT_ROWS 512         // ProbTable rows: expected identity i / 511 (analysis/ProbTable.cc:24)
#define SXC_PT_LEN 2048         // ProbTable columns: match length, clamped (analysis/ProbTable.cc:28, 79-80)

// One chunk: bytes live at seq_base + off, `len` bases (uppercase ASCII < 128).
typedef struct {
  int64_t off;
  int32_t len;
  int32_t flags;  // bit 0: contains a codec class with 1/3 entries (B, V, H, D); bit 1: only
                  // A, C, G, T; bit 2: only A, C, G, T, N; bits 8-31 (chunks with bit 0):
                  // row of the host-computed window entropy table (SXC_ENT_ROW floats per
                  // row: SXC_ENT_ROW / 2 windows of the forward strand, then of the reverse
                  // complement)
} sxc_chunk;

#define SXC_ENT_ROW 1024  // floats per entropy row: 2 strands x at most 512 windows

// One chunk pair of a batch (HomologyByXCorr::Align inputs plus the chunk metadata
// FilterMatches uses, analysis/HomologyByXCorrSlave.cc:168-253).
typedef struct {
  int32_t t_chunk;        // index into the target chunk table (bytes + resident spectra)
  int32_t q_chunk;        // index into the query chunk table
  int32_t t_chunk_start;  // SeqChunk::GetStart of the target chunk
  int32_t q_chunk_start;  // SeqChunk::GetStart of the query chunk
  int32_t t_record;       // target record id (t_result.target_id)
  int32_t q_record;       // query record id (t_result.query_id)
  int32_t q_record_len;   // ChunkManager::GetSize(query record) (t_result.query_size, RCQuery)
  int32_t fast;           // t_pair.fast: use cutoff_fast
  int32_t q_tail;         // query signal tail (index into the tail table), -1: none
} sxc_pair_dev;

// Signal tail: segments [first, first + n) of the segment table.
typedef struct {
  int32_t first;
  int32_t n;
} sxc_tail_dev;

// Tail segment: positions below `end` (and at or above the previous segment's end, or
// the chunk length) come from source signal `src` (4*N floats per source).
typedef struct {
  int32_t end;
  int32_t src;
} sxc_tail_seg_dev;

// Reference t_result (analysis/WorkQueue.h:23-33), 72 bytes on LP64.  Inside the library,
// a record whose device probability lies within the keep band of min_prob (keep_band) carries
// pad_[0] = SXC_REC_NEAR and its pair's index in the launch in pad_[1..4]; the host decides
// it and clears the pad before any record leaves the library.
#define SXC_REC_NEAR 1
typedef struct {
  uint64_t query_id;
  uint64_t target_id;
  uint64_t query_size;
  uint64_t qstart;
  uint64_t tstart;
  uint64_t len;
  uint8_t reverse;
  uint8_t pad_[7];
  double prob;
  double ident;
} sxc_result_dev;

// Twiddle tables: levels 3 .. 12 back to back, level L holding 2^(L-1) entries (a unit's lanes
// then read consecutive words of one level, not words 2^(12-L) apart in one table).
#define SXC_LUT_OFF(L) ((1 << ((L) - 1)) - 4)
#define SXC_LUT_LEN SXC_LUT_OFF(13)

// Constant tables built by the host library at context creation.
typedef struct {
  float lut[SXC_LUT_LEN];        // FFTReal's per-level cosine LUTs (FFTReal.hpp:352-374): level L
                                // (3..12) at SXC_LUT_OFF(L), (float)cos(i * PI / 2^L) for
                                // i < 2^(L-1); every level is a stride of level 12's
  unsigned char cls[128];       // byte -> codec class
  unsigned char rc[128];        // DNACodec::m_rc (DNAVector.cc:13-57)
  unsigned char score[SXC_NCLASS][SXC_NCLASS];  // (int)(100 * DNA_Equal + 0.5) per class pair
  unsigned char quarters[SXC_NCLASS][4];        // 4 * codec value when a multiple of 1/4
  double codec[SXC_NCLASS][4];  // DNA_A/C/G/T per class
  double equal[SXC_NCLASS][SXC_NCLASS];  // DNA_Equal per class pair (DNAVector.cc:390-403)
  double gc[SXC_NCLASS];        // DNA_C + DNA_G per class
  unsigned char rcls[128];      // byte -> codec class of its complement: cls[rc[byte] & 127]
  uint32_t qpack[2][128];       // strand, byte -> quarters of its class (strand 0) or of its
                                // complement's (strand 1), channel k in byte k
} sxc_tables;
