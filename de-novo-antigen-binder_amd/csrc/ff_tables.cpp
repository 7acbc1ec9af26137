// Force-field and geometry data; see ff_tables.h for the reference sources.
#include "ff_tables.h"

namespace dnab {

// Rmin/2 and epsilon per AtomType, src/amber.rs:46-63.
const NbParams kAmberNb[kNumAtomTypes] = {
    {1.9080f, 0.0860f},  // C
    {1.9080f, 0.0860f},  // CA
    {1.9080f, 0.0860f},  // CB
    {1.9080f, 0.0860f},  // CC
    {1.9080f, 0.1094f},  // CT
    {1.8240f, 0.1700f},  // N
    {1.8240f, 0.1700f},  // N2
    {1.8240f, 0.1700f},  // N3
    {1.8240f, 0.1700f},  // NA
    {1.8240f, 0.1700f},  // NB
    {1.6612f, 0.2100f},  // O
    {1.6612f, 0.2100f},  // O2
    {1.7210f, 0.2104f},  // OH
    {1.6837f, 0.1700f},  // OS
    {2.0000f, 0.2500f},  // S
    {2.0000f, 0.2500f},  // SH
};

// EEF1 (dg_ref, volume, lambda) per AtomType, src/amber.rs:414-431.
#define DNAB_EEF1_ROW(dg, vol, lam) {dg, vol, lam},
const Eef1Params kEef1[kNumAtomTypes] = {DNAB_EEF1_ROWS(DNAB_EEF1_ROW)};
#undef DNAB_EEF1_ROW

namespace {

// Heavy atoms per residue in topology order (backbone N, CA, C, O, CB, then
// side chain) with AMBER99SB charges and the hydrophobic flag,
// src/amber.rs:126-351.
const ResAtomDef kAtomsAla[] = {
    {" N  ", kN, -0.4157f, 0}, {" CA ", kCT, 0.0337f, 0}, {" C  ", kC, 0.5973f, 0},
    {" O  ", kO, -0.5679f, 0}, {" CB ", kCT, -0.1825f, 1}};
const ResAtomDef kAtomsArg[] = {
    {" N  ", kN, -0.3479f, 0},  {" CA ", kCT, -0.2637f, 0}, {" C  ", kC, 0.7341f, 0},
    {" O  ", kO, -0.5894f, 0},  {" CB ", kCT, -0.0007f, 1}, {" CG ", kCT, 0.0390f, 1},
    {" CD ", kCT, 0.0486f, 0},  {" NE ", kN2, -0.5295f, 0}, {" CZ ", kCA, 0.8076f, 0},
    {" NH1", kN2, -0.8627f, 0}, {" NH2", kN2, -0.8627f, 0}};
const ResAtomDef kAtomsAsn[] = {
    {" N  ", kN, -0.4157f, 0}, {" CA ", kCT, -0.0301f, 0}, {" C  ", kC, 0.5973f, 0},
    {" O  ", kO, -0.5679f, 0}, {" CB ", kCT, -0.2041f, 1}, {" CG ", kC, 0.7130f, 0},
    {" OD1", kO, -0.5931f, 0}, {" ND2", kN, -0.9191f, 0}};
const ResAtomDef kAtomsAsp[] = {
    {" N  ", kN, -0.5163f, 0},  {" CA ", kCT, 0.0381f, 0},  {" C  ", kC, 0.5366f, 0},
    {" O  ", kO, -0.5819f, 0},  {" CB ", kCT, -0.0303f, 1}, {" CG ", kC, 0.7994f, 0},
    {" OD1", kO2, -0.8014f, 0}, {" OD2", kO2, -0.8014f, 0}};
const ResAtomDef kAtomsCys[] = {
    {" N  ", kN, -0.4157f, 0}, {" CA ", kCT, 0.0213f, 0},  {" C  ", kC, 0.5973f, 0},
    {" O  ", kO, -0.5679f, 0}, {" CB ", kCT, -0.1231f, 1}, {" SG ", kSH, -0.3119f, 0}};
const ResAtomDef kAtomsGln[] = {
    {" N  ", kN, -0.4157f, 0}, {" CA ", kCT, -0.0031f, 0}, {" C  ", kC, 0.5973f, 0},
    {" O  ", kO, -0.5679f, 0}, {" CB ", kCT, -0.0036f, 1}, {" CG ", kCT, -0.0645f, 1},
    {" CD ", kC, 0.6951f, 0},  {" OE1", kO, -0.6086f, 0},  {" NE2", kN, -0.9407f, 0}};
const ResAtomDef kAtomsGlu[] = {
    {" N  ", kN, -0.5163f, 0}, {" CA ", kCT, 0.0397f, 0},  {" C  ", kC, 0.5366f, 0},
    {" O  ", kO, -0.5819f, 0}, {" CB ", kCT, 0.0560f, 1},  {" CG ", kCT, -0.0173f, 1},
    {" CD ", kC, 0.8054f, 0},  {" OE1", kO2, -0.8188f, 0}, {" OE2", kO2, -0.8188f, 0}};
const ResAtomDef kAtomsGly[] = {
    {" N  ", kN, -0.4157f, 0}, {" CA ", kCT, -0.0252f, 0}, {" C  ", kC, 0.5973f, 0},
    {" O  ", kO, -0.5679f, 0}};
const ResAtomDef kAtomsHis[] = {
    {" N  ", kN, -0.4157f, 0},  {" CA ", kCT, 0.0188f, 0},  {" C  ", kC, 0.5973f, 0},
    {" O  ", kO, -0.5679f, 0},  {" CB ", kCT, -0.0462f, 1}, {" CG ", kCC, -0.0266f, 0},
    {" ND1", kNB, -0.3811f, 0}, {" CD2", kCC, 0.1292f, 0},  {" CE1", kCC, 0.2057f, 0},
    {" NE2", kNA, -0.5727f, 0}};
const ResAtomDef kAtomsIle[] = {
    {" N  ", kN, -0.4157f, 0},  {" CA ", kCT, -0.0597f, 0}, {" C  ", kC, 0.5973f, 0},
    {" O  ", kO, -0.5679f, 0},  {" CB ", kCT, 0.1303f, 1},  {" CG1", kCT, -0.0430f, 1},
    {" CG2", kCT, -0.3204f, 1}, {" CD1", kCT, -0.0660f, 1}};
const ResAtomDef kAtomsLeu[] = {
    {" N  ", kN, -0.4157f, 0},  {" CA ", kCT, -0.0518f, 0}, {" C  ", kC, 0.5973f, 0},
    {" O  ", kO, -0.5679f, 0},  {" CB ", kCT, -0.1102f, 1}, {" CG ", kCT, 0.3531f, 1},
    {" CD1", kCT, -0.4121f, 1}, {" CD2", kCT, -0.4121f, 1}};
const ResAtomDef kAtomsLys[] = {
    {" N  ", kN, -0.3479f, 0}, {" CA ", kCT, -0.2400f, 0}, {" C  ", kC, 0.7341f, 0},
    {" O  ", kO, -0.5894f, 0}, {" CB ", kCT, -0.0094f, 1}, {" CG ", kCT, 0.0187f, 1},
    {" CD ", kCT, -0.0479f, 1}, {" CE ", kCT, -0.0143f, 1}, {" NZ ", kN3, -0.3854f, 0}};
const ResAtomDef kAtomsMet[] = {
    {" N  ", kN, -0.4157f, 0}, {" CA ", kCT, -0.0237f, 0}, {" C  ", kC, 0.5973f, 0},
    {" O  ", kO, -0.5679f, 0}, {" CB ", kCT, 0.0342f, 1},  {" CG ", kCT, 0.0018f, 1},
    {" SD ", kS, -0.2737f, 0}, {" CE ", kCT, -0.0536f, 1}};
const ResAtomDef kAtomsPhe[] = {
    {" N  ", kN, -0.4157f, 0},  {" CA ", kCT, -0.0024f, 0}, {" C  ", kC, 0.5973f, 0},
    {" O  ", kO, -0.5679f, 0},  {" CB ", kCT, -0.0343f, 1}, {" CG ", kCA, 0.0118f, 1},
    {" CD1", kCA, -0.1256f, 1}, {" CD2", kCA, -0.1256f, 1}, {" CE1", kCA, -0.1704f, 1},
    {" CE2", kCA, -0.1704f, 1}, {" CZ ", kCA, -0.1072f, 1}};
const ResAtomDef kAtomsPro[] = {
    {" N  ", kN, -0.2548f, 0}, {" CA ", kCT, -0.0266f, 0}, {" C  ", kC, 0.5896f, 0},
    {" O  ", kO, -0.5748f, 0}, {" CB ", kCT, -0.0070f, 1}, {" CG ", kCT, 0.0189f, 1},
    {" CD ", kCT, 0.0192f, 1}};
const ResAtomDef kAtomsSer[] = {
    {" N  ", kN, -0.4157f, 0}, {" CA ", kCT, -0.0249f, 0}, {" C  ", kC, 0.5973f, 0},
    {" O  ", kO, -0.5679f, 0}, {" CB ", kCT, 0.2117f, 0},  {" OG ", kOH, -0.6546f, 0}};
const ResAtomDef kAtomsThr[] = {
    {" N  ", kN, -0.4157f, 0}, {" CA ", kCT, 0.0764f, 0}, {" C  ", kC, 0.5973f, 0},
    {" O  ", kO, -0.5679f, 0}, {" CB ", kCT, 0.3654f, 0}, {" OG1", kOH, -0.6761f, 0},
    {" CG2", kCT, -0.2438f, 1}};
const ResAtomDef kAtomsTrp[] = {
    {" N  ", kN, -0.4157f, 0},  {" CA ", kCT, -0.0275f, 0}, {" C  ", kC, 0.5973f, 0},
    {" O  ", kO, -0.5679f, 0},  {" CB ", kCT, -0.0050f, 1}, {" CG ", kCB, -0.1415f, 1},
    {" CD1", kCB, -0.1638f, 1}, {" CD2", kCA, 0.1243f, 1},  {" NE1", kNA, -0.3418f, 0},
    {" CE2", kCA, 0.1380f, 1},  {" CE3", kCA, -0.2387f, 1}, {" CZ2", kCA, -0.2601f, 1},
    {" CZ3", kCA, -0.1972f, 1}, {" CH2", kCA, -0.1134f, 1}};
const ResAtomDef kAtomsTyr[] = {
    {" N  ", kN, -0.4157f, 0},  {" CA ", kCT, -0.0014f, 0}, {" C  ", kC, 0.5973f, 0},
    {" O  ", kO, -0.5679f, 0},  {" CB ", kCT, -0.0152f, 1}, {" CG ", kCA, -0.0011f, 1},
    {" CD1", kCA, -0.1906f, 1}, {" CD2", kCA, -0.1906f, 1}, {" CE1", kCA, -0.2341f, 1},
    {" CE2", kCA, -0.2341f, 1}, {" CZ ", kCA, 0.3226f, 1},  {" OH ", kOH, -0.5579f, 0}};
const ResAtomDef kAtomsVal[] = {
    {" N  ", kN, -0.4157f, 0},  {" CA ", kCT, 0.0145f, 0},  {" C  ", kC, 0.5973f, 0},
    {" O  ", kO, -0.5679f, 0},  {" CB ", kCT, 0.2985f, 1},  {" CG1", kCT, -0.3192f, 1},
    {" CG2", kCT, -0.3192f, 1}};

#define DNAB_N(a) static_cast<uint8_t>(sizeof(a) / sizeof((a)[0]))

// Top-5 rotamers per residue, src/rotamer.rs:33-160 (ALA, GLY have none).
const RotamerDef kRotArg[] = {{{-67, 180, 65, 85}, 0.22f},  {{-67, 180, 180, 85}, 0.18f},
                              {{-67, 180, 65, -85}, 0.12f}, {{-67, 180, 180, 180}, 0.10f},
                              {{62, 180, 65, 85}, 0.08f}};
const RotamerDef kRotAsn[] = {{{-65, -10, 0, 0}, 0.36f}, {{-65, 130, 0, 0}, 0.18f},
                              {{-174, 10, 0, 0}, 0.12f}, {{-174, -10, 0, 0}, 0.10f},
                              {{62, 10, 0, 0}, 0.08f}};
const RotamerDef kRotAsp[] = {{{-70, -15, 0, 0}, 0.38f}, {{-70, 170, 0, 0}, 0.22f},
                              {{-174, 25, 0, 0}, 0.16f}, {{62, 25, 0, 0}, 0.09f},
                              {{62, -15, 0, 0}, 0.06f}};
const RotamerDef kRotCys[] = {{{-65, 0, 0, 0}, 0.44f}, {{-179, 0, 0, 0}, 0.30f},
                              {{62, 0, 0, 0}, 0.21f},  {{62, 0, 0, 0}, 0.05f},
                              {{-65, 0, 0, 0}, 0.00f}};
const RotamerDef kRotGln[] = {{{-67, 180, 20, 0}, 0.25f}, {{-67, 180, -20, 0}, 0.18f},
                              {{-67, -65, 40, 0}, 0.12f}, {{-174, 65, -30, 0}, 0.10f},
                              {{62, 180, 20, 0}, 0.08f}};
const RotamerDef kRotGlu[] = {{{-67, 180, 20, 0}, 0.26f}, {{-67, 180, -20, 0}, 0.20f},
                              {{-67, -65, 40, 0}, 0.12f}, {{-174, 65, -20, 0}, 0.10f},
                              {{62, 180, -20, 0}, 0.08f}};
const RotamerDef kRotHis[] = {{{-62, -75, 0, 0}, 0.28f}, {{-62, 80, 0, 0}, 0.19f},
                              {{-174, -80, 0, 0}, 0.16f}, {{-174, 80, 0, 0}, 0.13f},
                              {{62, -80, 0, 0}, 0.08f}};
const RotamerDef kRotIle[] = {{{-60, 170, 0, 0}, 0.38f}, {{60, 170, 0, 0}, 0.20f},
                              {{-60, -60, 0, 0}, 0.17f}, {{-174, 170, 0, 0}, 0.12f},
                              {{60, -60, 0, 0}, 0.05f}};
const RotamerDef kRotLeu[] = {{{-60, 180, 0, 0}, 0.46f}, {{-60, 60, 0, 0}, 0.20f},
                              {{180, 60, 0, 0}, 0.14f},  {{60, 180, 0, 0}, 0.10f},
                              {{-174, 60, 0, 0}, 0.06f}};
const RotamerDef kRotLys[] = {{{-67, 180, 68, 180}, 0.21f},  {{-67, 180, 180, 180}, 0.17f},
                              {{-67, 180, -68, 180}, 0.12f}, {{-174, 180, 68, 180}, 0.09f},
                              {{62, 180, 68, 180}, 0.07f}};
const RotamerDef kRotMet[] = {{{-67, 180, 75, 0}, 0.24f}, {{-67, 180, -75, 0}, 0.18f},
                              {{-67, -65, 75, 0}, 0.12f}, {{-174, 65, 75, 0}, 0.10f},
                              {{62, 180, 75, 0}, 0.08f}};
const RotamerDef kRotPhe[] = {{{-65, 90, 0, 0}, 0.37f},  {{-65, -85, 0, 0}, 0.22f},
                              {{-174, 80, 0, 0}, 0.16f}, {{62, 90, 0, 0}, 0.12f},
                              {{-174, -80, 0, 0}, 0.06f}};
const RotamerDef kRotPro[] = {{{-65, 0, 0, 0}, 0.58f}, {{30, 0, 0, 0}, 0.28f},
                              {{-30, 0, 0, 0}, 0.14f}, {{-65, 0, 0, 0}, 0.00f},
                              {{-65, 0, 0, 0}, 0.00f}};
const RotamerDef kRotSer[] = {{{62, 0, 0, 0}, 0.37f}, {{-65, 0, 0, 0}, 0.35f},
                              {{-174, 0, 0, 0}, 0.23f}, {{62, 0, 0, 0}, 0.05f},
                              {{-65, 0, 0, 0}, 0.00f}};
const RotamerDef kRotThr[] = {{{62, 0, 0, 0}, 0.45f}, {{-60, 0, 0, 0}, 0.32f},
                              {{-174, 0, 0, 0}, 0.18f}, {{62, 0, 0, 0}, 0.05f},
                              {{-60, 0, 0, 0}, 0.00f}};
const RotamerDef kRotTrp[] = {{{-65, 96, 0, 0}, 0.32f},  {{-65, -84, 0, 0}, 0.20f},
                              {{-174, 96, 0, 0}, 0.15f}, {{62, 96, 0, 0}, 0.12f},
                              {{-174, -84, 0, 0}, 0.08f}};
const RotamerDef kRotTyr[] = {{{-65, 90, 0, 0}, 0.35f},  {{-65, -85, 0, 0}, 0.22f},
                              {{-174, 80, 0, 0}, 0.15f}, {{62, 90, 0, 0}, 0.12f},
                              {{-174, -80, 0, 0}, 0.06f}};
const RotamerDef kRotVal[] = {{{-60, 0, 0, 0}, 0.47f}, {{180, 0, 0, 0}, 0.36f},
                              {{60, 0, 0, 0}, 0.14f},  {{-60, 0, 0, 0}, 0.03f},
                              {{180, 0, 0, 0}, 0.00f}};

constexpr uint8_t X = 255;  // fixed dihedral
// Side-chain NERF recipes, src/rotamer.rs:288-406. Reference atom indices
// count into the residue's atom list: 0 N, 1 CA, 2 C, 3 O, 4 CB, 5.. side chain.
const BondDef kBdCys[] = {{0, 1, 4, 1.81f, 114.0f, 0.0f, 0}};
const BondDef kBdSer[] = {{0, 1, 4, 1.42f, 110.8f, 0.0f, 0}};
const BondDef kBdThr[] = {{0, 1, 4, 1.43f, 109.1f, 0.0f, 0}, {0, 1, 4, 1.52f, 111.5f, 120.0f, X}};
const BondDef kBdVal[] = {{0, 1, 4, 1.52f, 111.0f, 0.0f, 0}, {0, 1, 4, 1.52f, 111.0f, 120.0f, X}};
const BondDef kBdIle[] = {{0, 1, 4, 1.53f, 111.5f, 0.0f, 0},
                          {0, 1, 4, 1.52f, 111.5f, 120.0f, X},
                          {1, 4, 5, 1.52f, 114.0f, 0.0f, 1}};
const BondDef kBdLeu[] = {{0, 1, 4, 1.53f, 116.1f, 0.0f, 0},
                          {1, 4, 5, 1.52f, 111.0f, 0.0f, 1},
                          {1, 4, 5, 1.52f, 111.0f, 120.0f, X}};
const BondDef kBdMet[] = {{0, 1, 4, 1.53f, 114.1f, 0.0f, 0},
                          {1, 4, 5, 1.81f, 112.7f, 0.0f, 1},
                          {4, 5, 6, 1.79f, 100.9f, 0.0f, 2}};
const BondDef kBdPro[] = {{0, 1, 4, 1.53f, 104.0f, 0.0f, 0}, {1, 4, 5, 1.50f, 106.0f, 0.0f, X}};
const BondDef kBdPhe[] = {{0, 1, 4, 1.51f, 113.8f, 0.0f, 0},   {1, 4, 5, 1.39f, 120.7f, 0.0f, 1},
                          {1, 4, 5, 1.39f, 120.7f, 180.0f, X}, {4, 5, 6, 1.39f, 120.0f, 0.0f, X},
                          {4, 6, 7, 1.39f, 120.0f, 0.0f, X},   {5, 6, 8, 1.38f, 120.0f, 0.0f, X}};
const BondDef kBdTyr[] = {{0, 1, 4, 1.51f, 113.8f, 0.0f, 0},   {1, 4, 5, 1.39f, 120.7f, 0.0f, 1},
                          {1, 4, 5, 1.39f, 120.7f, 180.0f, X}, {4, 5, 6, 1.39f, 120.0f, 0.0f, X},
                          {4, 6, 7, 1.39f, 120.0f, 0.0f, X},   {5, 6, 8, 1.38f, 120.0f, 0.0f, X},
                          {6, 7, 9, 1.38f, 119.8f, 180.0f, X}};
const BondDef kBdTrp[] = {{0, 1, 4, 1.50f, 114.1f, 0.0f, 0},   {1, 4, 5, 1.37f, 126.8f, 0.0f, 1},
                          {1, 4, 5, 1.43f, 126.8f, 180.0f, X}, {4, 5, 6, 1.37f, 108.0f, 0.0f, X},
                          {4, 6, 7, 1.40f, 126.0f, 0.0f, X},   {6, 7, 8, 1.40f, 120.0f, 0.0f, X},
                          {7, 8, 9, 1.37f, 120.0f, 0.0f, X},   {8, 9, 10, 1.37f, 120.0f, 0.0f, X},
                          {9, 10, 11, 1.40f, 120.0f, 0.0f, X}};
const BondDef kBdHis[] = {{0, 1, 4, 1.50f, 113.8f, 0.0f, 0},   {1, 4, 5, 1.38f, 122.7f, 0.0f, 1},
                          {1, 4, 5, 1.36f, 131.5f, 180.0f, X}, {4, 5, 6, 1.32f, 108.0f, 0.0f, X},
                          {4, 6, 7, 1.36f, 108.0f, 0.0f, X}};
const BondDef kBdAsp[] = {{0, 1, 4, 1.52f, 113.4f, 0.0f, 0},
                          {1, 4, 5, 1.25f, 118.4f, 0.0f, 1},
                          {1, 4, 5, 1.25f, 118.4f, 180.0f, X}};
const BondDef kBdAsn[] = {{0, 1, 4, 1.52f, 112.6f, 0.0f, 0},
                          {1, 4, 5, 1.23f, 120.8f, 0.0f, 1},
                          {1, 4, 5, 1.33f, 116.5f, 180.0f, X}};
const BondDef kBdGlu[] = {{0, 1, 4, 1.53f, 114.1f, 0.0f, 0},
                          {1, 4, 5, 1.52f, 116.6f, 0.0f, 1},
                          {4, 5, 6, 1.25f, 118.4f, 0.0f, 2},
                          {4, 5, 6, 1.25f, 118.4f, 180.0f, X}};
const BondDef kBdGln[] = {{0, 1, 4, 1.53f, 114.1f, 0.0f, 0},
                          {1, 4, 5, 1.52f, 111.3f, 0.0f, 1},
                          {4, 5, 6, 1.23f, 120.8f, 0.0f, 2},
                          {4, 5, 6, 1.33f, 116.5f, 180.0f, X}};
const BondDef kBdLys[] = {{0, 1, 4, 1.53f, 114.1f, 0.0f, 0},
                          {1, 4, 5, 1.52f, 111.3f, 0.0f, 1},
                          {4, 5, 6, 1.52f, 111.3f, 0.0f, 2},
                          {5, 6, 7, 1.49f, 112.0f, 0.0f, 3}};
const BondDef kBdArg[] = {{0, 1, 4, 1.53f, 114.1f, 0.0f, 0}, {1, 4, 5, 1.52f, 111.3f, 0.0f, 1},
                          {4, 5, 6, 1.46f, 112.3f, 0.0f, 2}, {5, 6, 7, 1.33f, 124.2f, 0.0f, 3},
                          {6, 7, 8, 1.33f, 120.0f, 0.0f, X}, {6, 7, 8, 1.33f, 120.0f, 180.0f, X}};

}  // namespace

// Topology table indexed by AminoAcid, src/amber.rs:357-378 (ca_idx is 1
// for every residue; n_chi as listed there).
const ResTopology kResidueTopology[kNumAminoAcids] = {
    {kAtomsAla, DNAB_N(kAtomsAla), 1, 0}, {kAtomsArg, DNAB_N(kAtomsArg), 1, 4}, {kAtomsAsn, DNAB_N(kAtomsAsn), 1, 2},
    {kAtomsAsp, DNAB_N(kAtomsAsp), 1, 2}, {kAtomsCys, DNAB_N(kAtomsCys), 1, 1}, {kAtomsGln, DNAB_N(kAtomsGln), 1, 3},
    {kAtomsGlu, DNAB_N(kAtomsGlu), 1, 3}, {kAtomsGly, DNAB_N(kAtomsGly), 1, 0}, {kAtomsHis, DNAB_N(kAtomsHis), 1, 2},
    {kAtomsIle, DNAB_N(kAtomsIle), 1, 2}, {kAtomsLeu, DNAB_N(kAtomsLeu), 1, 2}, {kAtomsLys, DNAB_N(kAtomsLys), 1, 4},
    {kAtomsMet, DNAB_N(kAtomsMet), 1, 3}, {kAtomsPhe, DNAB_N(kAtomsPhe), 1, 2}, {kAtomsPro, DNAB_N(kAtomsPro), 1, 2},
    {kAtomsSer, DNAB_N(kAtomsSer), 1, 1}, {kAtomsThr, DNAB_N(kAtomsThr), 1, 1}, {kAtomsTrp, DNAB_N(kAtomsTrp), 1, 2},
    {kAtomsTyr, DNAB_N(kAtomsTyr), 1, 2}, {kAtomsVal, DNAB_N(kAtomsVal), 1, 1},
};

const RotamerDef* const kRotamerLib[kNumAminoAcids] = {
    nullptr, kRotArg, kRotAsn, kRotAsp, kRotCys, kRotGln, kRotGlu, nullptr, kRotHis, kRotIle,
    kRotLeu, kRotLys, kRotMet, kRotPhe, kRotPro, kRotSer, kRotThr, kRotTrp, kRotTyr, kRotVal};
const uint8_t kRotamerCount[kNumAminoAcids] = {0, 5, 5, 5, 5, 5, 5, 0, 5, 5,
                                               5, 5, 5, 5, 5, 5, 5, 5, 5, 5};

const BondDef* const kSidechainDefs[kNumAminoAcids] = {
    nullptr, kBdArg, kBdAsn, kBdAsp, kBdCys, kBdGln, kBdGlu, nullptr, kBdHis, kBdIle,
    kBdLeu,  kBdLys, kBdMet, kBdPhe, kBdPro, kBdSer, kBdThr, kBdTrp, kBdTyr, kBdVal};
const uint8_t kSidechainCount[kNumAminoAcids] = {
    0,            DNAB_N(kBdArg), DNAB_N(kBdAsn), DNAB_N(kBdAsp), DNAB_N(kBdCys),
    DNAB_N(kBdGln), DNAB_N(kBdGlu), 0,            DNAB_N(kBdHis), DNAB_N(kBdIle),
    DNAB_N(kBdLeu), DNAB_N(kBdLys), DNAB_N(kBdMet), DNAB_N(kBdPhe), DNAB_N(kBdPro),
    DNAB_N(kBdSer), DNAB_N(kBdThr), DNAB_N(kBdTrp), DNAB_N(kBdTyr), DNAB_N(kBdVal)};

const char kAminoOneLetter[kNumAminoAcids + 1] = "ARNDCQEGHILKMFPSTWYV";

// src/atom.rs:55-62
float aa_charge(int aa) {
  switch (aa) {
    case kArg: case kLys: return 1.0f;
    case kAsp: case kGlu: return -1.0f;
    case kHis: return 0.1f;
    default: return 0.0f;
  }
}

// src/atom.rs:66-76
float aa_lj_epsilon(int aa) {
  switch (aa) {
    case kGly: return 0.05f;
    case kAla: return 0.10f;
    case kVal: case kIle: case kLeu: return 0.15f;
    case kPhe: case kTrp: case kTyr: return 0.20f;
    case kArg: case kLys: return 0.12f;
    case kAsp: case kGlu: return 0.13f;
    default: return 0.11f;
  }
}

// src/atom.rs:80-90
float aa_lj_sigma(int aa) {
  switch (aa) {
    case kGly: return 2.5f;
    case kAla: return 3.0f;
    case kVal: return 3.5f;
    case kIle: case kLeu: return 3.8f;
    case kPhe: case kTrp: case kTyr: return 4.2f;
    case kArg: case kLys: return 4.0f;
    default: return 3.2f;
  }
}

// src/atom.rs:93-104
uint8_t aa_is_hydrophobic(int aa) {
  switch (aa) {
    case kAla: case kVal: case kIle: case kLeu: case kMet: case kPhe: case kTrp: return 1;
    default: return 0;
  }
}

}  // namespace dnab
