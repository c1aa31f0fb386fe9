// Hybrid near/far field: cell grid, moments, far-field operator, launches.
#pragma once
#include "kernels.hpp"
#include "pair_special.hpp"

#include <vector>

namespace ddd {

constexpr int kMoments = 90;  // cell features: orders 0, 1, 2 of b (x) t L
constexpr int kPoly = 10;     // local basis: 1, x, y, z, x^2, y^2, z^2, xy, xz, yz
constexpr int kLocal = 6 * kPoly;          // stress components x basis
constexpr int kChan = kLocal * kMoments;   // entries of one W(o)
// Real channels travel through the FFTs two per complex channel.
constexpr int kMomPairs = kMoments / 2;
constexpr int kLocPairs = kLocal / 2;
constexpr int kChanPairs = kChan / 2;

struct GridDev {
  double lo[3];
  double h;
  int G;
};

int fft_size_for(int G);
void launch_cells(const double* seg, int64_t ld, int n, const GridDev& g, int* cellOf,
                  unsigned* counts, unsigned* cursor, int* start, int* tmp, int* order,
                  hipStream_t st);
void launch_moments(const double* seg, int64_t ld, const GridDev& g, const int* start,
                    const int* order, float* M, hipStream_t st);
void launch_far_field(const float* M, const GridDev& g, const float2* What, const float2* tw,
                      float2* Mc, float2* Lc, float* L, hipStream_t st);
// What: [half frequency][90 moments][60 locals] complex, f_z in [0, NP/2] (NP^2 (NP/2 + 1)
// frequencies); scratch: kChanPairs NP^3 complex; Mc: kMomPairs NP^3;
// Lc: kLocPairs NP^3.
void launch_operator_spectrum(const float* W, const GridDev& g, int rn, const float2* tw,
                              float2* scratch, float2* What, hipStream_t st);
void launch_far_apply(const double* seg, int64_t ld, int n, const GridDev& g, const int* cellOf,
                      const float* L, double* out, hipStream_t st);
// Device arrays of the near-field tile list (see farfield.hip).
struct NearListDev {
  // inputs: cell binning
  const int* cellOf;
  const unsigned* counts;
  const int* start;
  const int* order;
  const int* mrank;     // Morton rank of each cell
  const int* rankCell;  // cell of each Morton rank
  // outputs
  unsigned* mcount;
  int* mstart;          // nc + 1
  int64_t ldS;          // pad64(numSeg): sorted copy leading dimension
  double* segS;         // F_COUNT x ldS
  int* key;             // packed cell coordinates per sorted segment
  int* perm;            // original index per sorted segment (-1: padding)
  int* box;             // 6 per block
  unsigned* rowCount;
  int* rowStart;        // nbS + 1; rowStart[nbS] = number of tiles
  unsigned* colCount;
  int* colStart;        // nbS + 1
  int2* list;           // (I, J) per tile
  int* tileIdx;         // nbS x nbS, J >= I
  int* colList;
  int rn;
};

void morton_ranks(int G, std::vector<int>& mrank, std::vector<int>& rankCell);
// Sorted copy, block boxes, per-row tile counts and rowStart (its last entry
// is the number of tiles); then, with list and colList sized, the lists.
void launch_near_list_count(const double* seg, int64_t ld, int n, const GridDev& g, int rn,
                            NearListDev& L, hipStream_t st);
void launch_near_list_fill(const NearListDev& L, hipStream_t st);
void launch_near_reduce(const NearListDev& L, int nc, const double* rowPart, const double* colPart,
                        double* out, hipStream_t st);
// Near-field tiles: workgroups of the main kernel and the staging entries
// (uint16) each one owns (two per slot of a tile); stage holds
// kNearGrid * kNearStage entries.  The main kernel lists the near-parallel
// pairs (spList, capacity spCap; specOff / specCnt per tile) and counts them
// in counters[3]; counters (four) are zeroed before the call.
constexpr int kNearGrid = 2048;
constexpr int kNearStage = 2 * kTile * kTile;
void launch_near_tiles(const double* segS, int64_t ldS, const int* key, int rn, const int2* list,
                       int nTiles, const SegK<double>& K, const MixedRule& mr, double* rowPart,
                       double* colPart, unsigned short* stage, unsigned long long* spList,
                       long long spCap, long long* specOff, int* specCnt,
                       unsigned long long* counters, hipStream_t st);
// The listed near-parallel pairs (total <= spCap): evaluated into res
// (12 x spCap doubles) and added to their tiles' partial sums.
void launch_near_special(const double* segS, int64_t ldS, const int2* list, int nTiles,
                         const SegK<double>& K, const unsigned long long* spList, long long total,
                         long long spCap, const long long* specOff, const int* specCnt, double* res,
                         double* rowPart, double* colPart, unsigned* flags, hipStream_t st);

}  // namespace ddd
