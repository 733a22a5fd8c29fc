// layout.hpp -- host-side description of the device-resident operator layout.
//
// An operator lives in one contiguous FP64 device buffer:
//   [E0 | pad to 32 doubles][F: nsp x nsp row-major one-body matrix][V]
// V holds, for every (J,P,Tz) channel block q, the FULL dim x dim row-major
// matrix over that block's pair basis, pairs ordered [pp | ph | hh] exactly as
// the reference lists them (allocate_blocks, basic_IMSRG.f90:477-641).  The
// reference keeps only six upper sub-blocks (gam1..gam6) and implies the rest
// through herm; keeping the full square block makes every 222 contraction a
// plain GEMM and every norm a flat sum.  Lower triangles are always
// herm * upper (diagonal zero when herm = -1).
//
// Pandya (cross-coupled) matrices and the ph product W live in context
// workspaces with the reference's block order (cross_coupled.f90:61-177).
// All angular momenta are doubled; orbits are 0-based.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace imsrg {

struct Layout {
  // orbits
  int nsp = 0, nhole = 0, npart = 0;
  std::vector<int> n, l, jj, tz, con, holes, parts;
  int jtot_max = 0, lmax = 0;
  // (l, 2j, 2tz) groups, members in orbit order
  std::vector<int> grp_of, grp_start, grp_cnt, grp_list;

  // two-body (J,P,Tz) blocks: J slowest, Tz = -1..1, P fastest
  int nblocks = 0;
  std::vector<int> bJ, bP, bTz, bnp, bnb, bnh, bdim;
  std::vector<int64_t> boff;    // element offset of block q inside V
  int64_t vsize = 0;            // elements of V (sum dim^2, blocks 32-aligned)
  std::vector<int> pbase;       // first global pair index of block q
  int npairs = 0;               // sum of dims
  std::vector<int> pair_a, pair_b, pair_q;  // per global pair (a <= b)
  // pidx[(J/2)*nsp*nsp + a*nsp + b] (a <= b): position inside block's full matrix
  std::vector<int> pidx;

  // Hartree-Fock tables.  hf_gorder: the (l,j,tz) groups in the reference's
  // single-particle block order (2tz, l, 2j; allocate_sp_mat, basic_IMSRG.f90:
  // 2600-2660); hf_b/hf_d: every ordered orbit pair of one group, groups in
  // that order; hf_a/hf_c: the pairs with a <= c.  hf_nb[hf_nb_start[r] ..]:
  // for pair row r = (i,j) of block q, the rows (a,b) of q with a, b in the
  // groups of i, j (either order): the non-zeros of column r of the pair-basis
  // transform (transform_2b_to_HF, HF_mod.f90:494-619), positions in the block.
  std::vector<int> hf_gorder, hf_b, hf_d, hf_a, hf_c, hf_nb_start, hf_nb;

  // one-body "upper" outputs: (p,q) with p<=q inside hh/pp, all (particle,hole)
  std::vector<int> ob_p, ob_q;

  // 122 neighbours: for global pair r=(a,b), entries replacing a or b by an
  // orbit of the same (l,j,tz) group.  nb_pos is the new pair's position in
  // the same block; coef = exchange phase * norm(new)/norm(old).
  std::vector<int> nb_start;    // npairs + 1
  std::vector<int> nb_pos, nb_x, nb_i;
  std::vector<double> nb_coef;
  // 122 row sets: the rows of one block whose two orbits come from one
  // unordered pair of (l,j,tz) groups (the neighbours of every row of a set
  // lie in the set).  Set s is rows s122_rows[s122_start[s] .. s122_start[s+1])
  // (global pair indices) of block s122_q[s]; nb_loc is a neighbour entry's
  // index inside its row's set.  Sets of at most 32 rows with at most two
  // entries per slot (row k, neighbour row k') run in the set kernel: their
  // slots, row-major, start at s122_ab[s] / 2 (-1 for the other sets), slot i
  // holding f-matrix indices s122_sx[2i], [2i+1] and coefficients s122_sc
  // (absent entries: index 0, coefficient 0); work items (set, 64-column
  // chunk) as pairs, larger sets first.  The rows of the other sets go to
  // the row kernel (s122_frows).
  std::vector<int> s122_q, s122_start, s122_rows, nb_loc, s122_frows, s122_items, s122_sx;
  std::vector<long long> s122_ab;
  std::vector<double> s122_sc;

  // Pandya blocks: JC fastest, |Tz|, parity slowest
  int ncc = 0;
  std::vector<int> ccJ, ccrlen, ccnph, ccTz, ccP;
  std::vector<int64_t> ccoff;   // element offset of CC block (rlen x nph, column-major)
  std::vector<int64_t> cwoff;   // element offset of W block (rlen x rlen, row-major)
  int64_t ccsize = 0, cwsize = 0;
  std::vector<int> ccrow;       // [(q*nsp + a)*nsp + b] -> row, -1
  std::vector<int> ccrbase;     // first global row of block q
  std::vector<int> ccrow_a, ccrow_b;  // per global row (-1 for the zero row that makes the count even)
  // Rows of a Pandya block come in pairs: (a,b) at 2u and (b,a) at 2u+1 (a < b),
  // then the (a,a) rows, then a zero row if needed, so ccrlen is even.  The reference's
  // row order (all valid ordered pairs, first orbit major) is kept for copies
  // to the host: ccref[ccrefbase[q] + k] = device row of reference row k.
  std::vector<int> ccref, ccrefbase, ccrlen_ref;
  std::vector<int> cccbase;     // first global column of block q
  std::vector<int> cccol_p, cccol_h;  // per global column (particle, hole)

  // (i,j,k,l) quadruples owning at least one upper element (ij,kl;J), packed
  // i<<24|j<<16|k<<8|l and sorted by (i,k,j,l) for the ph recoupling kernel
  std::vector<uint32_t> ph_quads;
  std::vector<uint32_t> ph_terms;   // quad<<6 | side<<5 | m (side 0: T1, 1: T2), or kPhPadTerm
  std::vector<uint32_t> ph_toff;    // per upper element: offset of its quadruple's T1 terms
  std::vector<int64_t> ph_ubase;    // per global pair row: first upper element
  std::vector<int64_t> ph_tbase;    // per global pair row: the smallest ph_toff of its upper elements
  std::vector<uint64_t> ph_desc;    // per upper element: T offset, run lengths, 6j runs, flags
  std::vector<uint32_t> ph_zblk;    // per J' term: 16-byte half of a 2x2 Z block (units of 2 doubles)
  std::vector<uint8_t> ph_zsel;     // per J' term: element selector / flags
  std::vector<uint32_t> ph_two;     // the J' terms whose two Z elements lie in two halves of a block

  // dense 6j table {j1 j2 J3; j4 j5 J6}: j1,j2,j4,j5 half-integer, J3,J6 integer,
  // index order [j1][j2][J3][j4][j5][J6] so that the (j4,j5,J6) a wavefront
  // varies over sit in one small contiguous slab
  int nhalf = 0, nJ = 0;
  std::vector<double> sixj;
  std::vector<double> ap6_val;      // packed non-zero 6j runs per ph-apply row key (ji, jj, J)
  std::vector<int> ap6_base, ap6_off;
  int ap6_max = 0;

  int64_t neq = 0;              // length of the reference vectorize stream
  int64_t f_off = 32;           // F offset inside the operator buffer
  int64_t v_off = 0;            // V offset inside the operator buffer
  int64_t op_size = 0;          // doubles per operator buffer

  // algorithmic work counts (per commutator)
  double flop_pphh = 0, flop_ph = 0;

  std::string build(int nsp_, const int *n_, const int *l_, const int *jj_, const int *tz_,
                    const int *con_);

  // host conversions between the reference vectorize stream and the full layout
  // gradient with respect to the vectorize stream of a full-layout cotangent:
  // the transpose of unpack_vec (every full element unpack_vec writes adds
  // its cotangent, times the factor it was written with, to its source)
  void unpack_vec_transpose(const double *cot, int herm, double *v) const;
  void unpack_vec(const double *vec, int herm, double *opbuf) const;   // repackage
  void pack_vec(const double *opbuf, double *vec) const;               // vectorize

  int block_of(int J, int a, int b) const {  // block_index, basic_IMSRG.f90:1064-1070
    int P = (l[a] + l[b]) % 2, T = (tz[a] + tz[b]) / 2;
    return 3 * J + 2 * (T + 1) + P;  // J doubled
  }
};

double racah6j(int j1, int j2, int j3, int j4, int j5, int j6);  // doubled arguments

// Interaction files (interaction_IO.cpp): the reference's H%hospace, H%com_hw,
// H%lawson_beta, Aneut + Aprot and ham_type, and for me2j files the file's own
// single-particle basis cuts.
struct ReadParams {
  double hw = 0.0, com_hw = 0.0, lawson_beta = 0.0;
  int A = 0, htype = 0, emax_file = 0, lmax_file = 0;
};
// Sets the matrix elements the file lists in `img`, the host image of one
// operator in this layout; format 0 ascii, 1 gz, 2 me2j, 3 me2b (path2 = its
// one-body .me1b file).  0 or -1 with err.
int read_interaction_file(const Layout &L, int format, const char *path, const char *path2, const ReadParams &p,
                          double *img, std::string &err);

// a zero J' term that pads a quadruple's run to an even length
constexpr uint32_t kPhPadTerm = 0xFFFFFFFFu;

}  // namespace imsrg
