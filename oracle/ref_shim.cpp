// C-callable entry points into the reference's own CPU hashers, compiled
// together with the reference sources (native/poseidon/poseidon.cpp,
// native/poseidon2/poseidon2.cu, native/keccak/keccak.cu,
// native/monolith/monolith.cu, native/poseidon/poseidon_bn128.cu,
// native/merkle/merkle.cpp) by oracle/Makefile
// into oracle/_ref/libzkref.so.  Test infrastructure only: used to check the
// oracle restatement against the reference itself in the build container.
#include "types/int_types.h"
#include "poseidon/poseidon.hpp"
#include "poseidon2/poseidon2.hpp"
#include "keccak/keccak.hpp"
#include "monolith/monolith.hpp"
#include "poseidon/poseidon_bn128.hpp"

extern "C" void ref_hash_one(u64 *in, u64 n, u64 *out, int hash_type)
{
    if (hash_type == 3) Poseidon2Hasher::cpu_hash_one(in, n, out);
    else if (hash_type == 1) KeccakHasher::cpu_hash_one(in, n, out);
    else if (hash_type == 4) MonolithHasher::cpu_hash_one(in, n, out);
    else if (hash_type == 2) PoseidonBN128Hasher::cpu_hash_one(in, n, out);
    else PoseidonHasher::cpu_hash_one(in, n, out);
}

extern "C" void ref_hash_two(u64 *l, u64 *r, u64 *out, int hash_type)
{
    if (hash_type == 3) Poseidon2Hasher::cpu_hash_two(l, r, out);
    else if (hash_type == 1) KeccakHasher::cpu_hash_two(l, r, out);
    else if (hash_type == 4) MonolithHasher::cpu_hash_two(l, r, out);
    else if (hash_type == 2) PoseidonBN128Hasher::cpu_hash_two(l, r, out);
    else PoseidonHasher::cpu_hash_two(l, r, out);
}
