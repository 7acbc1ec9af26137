// Parameters of the H-bond, EEF1 solvation and disulfide terms, shared by the
// host tables (ff_tables.cpp) and the energy kernels (pair.hip) so that both
// read one copy of the numbers.
//   H-bond Gaussian well          src/energy.rs:23-39
//   H-bond donor / acceptor types src/amber.rs:65-84
//   EEF1 table                    src/amber.rs:396-431, src/solvation.rs:21-26
//   disulfide Gaussian well       src/allatom.rs:155-180
#pragma once
#include <cstdint>

namespace dnab {

constexpr float kHbR0 = 2.9f;         // A, ideal donor-acceptor distance
constexpr float kHbSigma = 0.35f;     // A, well width
constexpr float kHbDepth = 1.5f;      // kcal/mol
constexpr float kHbCutoffSq = 25.0f;  // A^2
constexpr float kPi32 = 5.568328f;    // pi^(3/2), as the reference writes it
constexpr float kSolvCutoffSq = 81.0f;  // (9 A)^2
constexpr float kSsD0 = 2.05f;        // A, ideal S-S bond
constexpr float kSsSigma = 0.4f;
constexpr float kSsDepth = 4.0f;      // kcal/mol
constexpr float kSsCutoff = 4.0f;     // A

// AtomType bit masks (C, CA, CB, CC, CT, N, N2, N3, NA, NB, O, O2, OH, OS, S,
// SH = bits 0..15): donors N, N2, N3, NA, OH, SH; acceptors NB, O, O2, OH, OS.
constexpr uint32_t kHbondDonorMask = (1u << 5) | (1u << 6) | (1u << 7) | (1u << 8) | (1u << 12) |
                                     (1u << 15);
constexpr uint32_t kHbondAcceptorMask = (1u << 9) | (1u << 10) | (1u << 11) | (1u << 12) |
                                        (1u << 13);

// EEF1 (dg_ref kcal/mol, volume A^3, lambda A) per AtomType, in type order.
#define DNAB_EEF1_ROWS(X)   \
  X(0.000f, 14.7f, 3.50f)   \
  X(-0.187f, 11.7f, 3.70f)  \
  X(-0.187f, 11.7f, 3.70f)  \
  X(-0.187f, 11.7f, 3.70f)  \
  X(-0.187f, 11.7f, 3.70f)  \
  X(-6.700f, 0.0f, 3.50f)   \
  X(-10.000f, 15.0f, 3.50f) \
  X(-20.000f, 15.0f, 3.50f) \
  X(-6.700f, 15.0f, 3.50f)  \
  X(-6.700f, 15.0f, 3.50f)  \
  X(-5.330f, 14.2f, 3.15f)  \
  X(-10.000f, 14.2f, 3.15f) \
  X(-5.920f, 12.5f, 3.15f)  \
  X(-5.920f, 12.5f, 3.15f)  \
  X(-3.240f, 32.9f, 3.90f)  \
  X(-3.240f, 32.9f, 3.90f)

}  // namespace dnab
