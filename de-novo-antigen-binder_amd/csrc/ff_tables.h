// Force-field and geometry tables used to build atom clouds on the host.
//
// The numbers are the reference's force-field data, kept value-for-value so
// that atom clouds built here carry exactly the parameters the reference's
// energy code sees:
//   AMBER99SB non-bonded table        src/amber.rs:46-63
//   per-residue heavy-atom topologies src/amber.rs:126-378
//   backbone-independent rotamers     src/rotamer.rs:33-184
//   side-chain NERF build recipes     src/rotamer.rs:288-430
//   residue-level (Calpha) parameters src/atom.rs:55-104
//   EEF1 solvation parameters         src/amber.rs:396-431 (numbers in ff_terms.h)
#pragma once
#include <cstdint>

#include "ff_terms.h"

namespace dnab {

constexpr int kNumAtomTypes = 16;   // src/amber.rs:32
constexpr int kNumAminoAcids = 20;  // src/atom.rs:27

// AtomType order of src/amber.rs:13-30.
enum AtomType : uint8_t {
  kC = 0, kCA, kCB, kCC, kCT, kN, kN2, kN3, kNA, kNB, kO, kO2, kOH, kOS, kS, kSH
};

// AminoAcid order of src/atom.rs:4-25.
enum AminoAcid : uint8_t {
  kAla = 0, kArg, kAsn, kAsp, kCys, kGln, kGlu, kGly, kHis, kIle,
  kLeu, kLys, kMet, kPhe, kPro, kSer, kThr, kTrp, kTyr, kVal
};

struct NbParams {
  float r_min_half;  // Rmin/2, Angstrom
  float epsilon;     // kcal/mol
};

struct ResAtomDef {
  const char* name;     // PDB-style 4-char name
  uint8_t atom_type;    // AtomType
  float charge;         // elementary charges
  uint8_t hydrophobic;  // 1 = counts for the hydrophobic pair bonus
};

struct ResTopology {
  const ResAtomDef* atoms;
  uint8_t n_atoms;
  uint8_t ca_idx;
  uint8_t n_chi;
};

struct RotamerDef {
  float chi[4];  // degrees
  float probability;
};

// One NERF placement: new atom D from reference atoms (a, b, c) of the
// residue's growing atom list. chi_idx 0..3 takes the dihedral from chi,
// 255 takes dihedral_fixed.
struct BondDef {
  uint8_t ref_a, ref_b, ref_c;
  float bond_len;
  float angle_deg;
  float dihedral_fixed;
  uint8_t chi_idx;
};

// EEF1 parameters per AtomType (Lazaridis & Karplus 1999), src/amber.rs:406-431.
struct Eef1Params {
  float dg_ref;  // reference solvation free energy, kcal/mol
  float vol;     // volume, A^3
  float lambda;  // correlation length, A
};

extern const NbParams kAmberNb[kNumAtomTypes];
extern const Eef1Params kEef1[kNumAtomTypes];
// Topology slot of the Cys SG atom (the disulfide term, src/allatom.rs:139-150).
constexpr int kCysSgSlot = 5;
extern const ResTopology kResidueTopology[kNumAminoAcids];
extern const RotamerDef* const kRotamerLib[kNumAminoAcids];
extern const uint8_t kRotamerCount[kNumAminoAcids];
extern const BondDef* const kSidechainDefs[kNumAminoAcids];
extern const uint8_t kSidechainCount[kNumAminoAcids];
extern const char kAminoOneLetter[kNumAminoAcids + 1];

// Residue-level (Calpha) parameters, src/atom.rs:55-104.
float aa_charge(int aa);
float aa_lj_epsilon(int aa);
float aa_lj_sigma(int aa);
uint8_t aa_is_hydrophobic(int aa);

}  // namespace dnab
