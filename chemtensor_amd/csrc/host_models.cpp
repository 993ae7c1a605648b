// Inputs of the DMRG path: chemtensor's random numbers (PCG32 pairs, Box-Muller, Floyd
// sampling; src/util/rng.c:15-125 on the PCG algorithm of src/util/pcg_basic.c), the seeded
// random MPS of src/state/mps.c:93-178 and the XXZ / Fermi-Hubbard MPOs whose operators are
// those of src/operator/hamiltonian.c:102-155 and :240-366 (built here directly as
// finite-state-machine MPOs; the virtual bond basis differs from the reference's graph
// construction, the operator is identical).

#include <cmath>
#include <cstring>
#include <vector>

#include "ct_internal.h"

namespace ctg {

static uint32_t pcg_next(pcg32_random* r)
{
	const uint64_t old = r->state;
	r->state = old * 6364136223846793005ULL + r->inc;
	const uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
	const uint32_t rot = (uint32_t)(old >> 59u);
	return (xs >> rot) | (xs << ((-rot) & 31));
}

static void pcg_seed(uint64_t initstate, uint64_t initseq, pcg32_random* r)
{
	r->state = 0u;
	r->inc = (initseq << 1u) | 1u;
	pcg_next(r);
	r->state += initstate;
	pcg_next(r);
}

static uint64_t rand_u64(rng_state* s)
{
	const uint64_t hi = pcg_next(&s->gen[0]);
	return (hi << 32) | pcg_next(&s->gen[1]);
}

static uint64_t rand_bounded(uint64_t bound, rng_state* s)
{
	const uint64_t threshold = (0 - bound) % bound;
	for (;;) {
		const uint64_t r = rand_u64(s);
		if (r >= threshold) return r % bound;
	}
}

static void rand_choice(uint64_t bound, uint64_t num, rng_state* s, uint64_t* ret)
{
	for (uint64_t i = 0; i < num; i++) {
		const uint64_t j = bound - num + 1 + i;
		const uint64_t t = rand_bounded(j, s);
		bool drawn = false;
		for (uint64_t k = 0; k < i; k++) if (ret[k] == t) { drawn = true; break; }
		ret[i] = drawn ? j - 1 : t;
	}
}

}  // namespace ctg

using namespace ctg;

extern "C" {

void seed_rng_state(const uint64_t seed, struct rng_state* state)
{
	pcg32_random tmp;
	pcg_seed(seed, 6364136223846793005ULL * seed + 1442695040888963407ULL, &tmp);
	uint64_t w[8];
	for (int i = 0; i < 8; i++) w[i] = pcg_next(&tmp);
	const uint64_t seed1 = (w[0] << 32) | w[1], seed2 = (w[2] << 32) | w[3];
	const uint64_t seq1 = (w[4] << 32) | w[5];
	uint64_t seq2 = (w[6] << 32) | w[7];
	const uint64_t mask = ~0ull >> 1;
	if ((seq1 & mask) == (seq2 & mask)) seq2 = ~seq2;
	pcg_seed(seed1, seq1, &state->gen[0]);
	pcg_seed(seed2, seq2, &state->gen[1]);
}

double randn(struct rng_state* state)
{
	double u1 = std::ldexp((double)rand_u64(state), -64);
	const double u2 = std::ldexp((double)rand_u64(state), -64);
	if (u1 == 0) u1 = 1;
	return std::sqrt(-2 * std::log(u1)) * std::sin(6.2831853071795864769 * u2);
}

void ct_fill_random_normal(struct block_sparse_tensor* t, double alpha, struct rng_state* rng)
{
	ct_long nb = 1;
	for (int i = 0; i < t->ndim; i++) nb *= t->dim_blocks[i];
	for (ct_long k = 0; k < nb; k++) {
		dense_tensor* b = t->blocks[k];
		if (!b) continue;
		ct_long n = 1;
		for (int i = 0; i < b->ndim; i++) n *= b->dim[i];
		double* d = (double*)b->data;
		for (ct_long j = 0; j < n; j++) d[j] = alpha * randn(rng) + 0.0;
	}
}

static void random_bond_qnums(int nsites, ct_long d, const qnumber* qsite, qnumber sector, ct_long max_vdim, rng_state* rng,
	std::vector<std::vector<qnumber>>& qb)
{
	qb.assign((size_t)nsites + 1, {});
	qb[0] = { 0 };
	qb[nsites] = { sector };
	for (int l = 1; l < (nsites + 1) / 2; l++) {
		std::vector<qnumber> full;
		for (qnumber a : qb[l - 1]) for (ct_long j = 0; j < d; j++) full.push_back(a + qsite[j]);
		if ((ct_long)full.size() <= max_vdim) qb[l] = full;
		else {
			std::vector<uint64_t> idx((size_t)max_vdim);
			rand_choice(full.size(), (uint64_t)max_vdim, rng, idx.data());
			for (uint64_t i : idx) qb[l].push_back(full[i]);
		}
	}
	for (int l = nsites - 1; l >= (nsites + 1) / 2; l--) {
		std::vector<qnumber> full;
		for (qnumber a : qb[l + 1]) for (ct_long j = 0; j < d; j++) full.push_back(a - qsite[j]);
		if ((ct_long)full.size() <= max_vdim) qb[l] = full;
		else {
			std::vector<uint64_t> idx((size_t)max_vdim);
			rand_choice(full.size(), (uint64_t)max_vdim, rng, idx.data());
			for (uint64_t i : idx) qb[l].push_back(full[i]);
		}
	}
}

void construct_random_mps(const enum numeric_type dtype, const int nsites, const ct_long d, const qnumber* qsite, const qnumber qnum_sector, const ct_long max_vdim, struct rng_state* rng_state, struct mps* mps)
{
	std::vector<std::vector<qnumber>> qb;
	random_bond_qnums(nsites, d, qsite, qnum_sector, max_vdim, rng_state, qb);
	mps->nsites = nsites;
	mps->d = d;
	mps->qsite = (qnumber*)aligned_alloc(16, ((size_t)d * sizeof(qnumber) + 15) / 16 * 16);
	memcpy(mps->qsite, qsite, (size_t)d * sizeof(qnumber));
	mps->a = (block_sparse_tensor*)calloc((size_t)nsites, sizeof(block_sparse_tensor));
	for (int i = 0; i < nsites; i++) {
		const ct_long dim[3] = { (ct_long)qb[i].size(), d, (ct_long)qb[i + 1].size() };
		const tensor_axis_direction ad[3] = { TENSOR_AXIS_OUT, TENSOR_AXIS_OUT, TENSOR_AXIS_IN };
		const qnumber* qn[3] = { qb[i].data(), qsite, qb[i + 1].data() };
		allocate_block_sparse_tensor(dtype, 3, dim, ad, qn, &mps->a[i]);
	}
	for (int i = 0; i < nsites; i++) {
		const ct_long nelem = mps->a[i].dim_logical[0] * mps->a[i].dim_logical[1] * mps->a[i].dim_logical[2];
		ct_fill_random_normal(&mps->a[i], 1.0 / std::sqrt((double)nelem), rng_state);
	}
}

void delete_mps(struct mps* mps)
{
	for (int i = 0; i < mps->nsites; i++) delete_block_sparse_tensor(&mps->a[i]);
	free(mps->a);
	mps->a = nullptr;
	free(mps->qsite);
	mps->qsite = nullptr;
	mps->nsites = 0;
	mps->d = 0;
}

void delete_mpo(struct mpo* mpo)
{
	for (int i = 0; i < mpo->nsites; i++) delete_block_sparse_tensor(&mpo->a[i]);
	free(mpo->a);
	mpo->a = nullptr;
	free(mpo->qsite);
	mpo->qsite = nullptr;
	mpo->nsites = 0;
	mpo->d = 0;
}

}  // extern "C"

namespace ctg {

// MPO of a finite-state machine: bulk tensor W[ml][p'][p][mr] (nb states, start 0, final nb-1);
// the first site keeps row 0, the last site column nb-1 (mpo.c:59-175 layout and directions)
static void mpo_from_fsm(int nsites, ct_long d, const qnumber* qsite, int nb, const qnumber* qbond, const std::vector<double>& w, mpo* out)
{
	out->nsites = nsites;
	out->d = d;
	out->qsite = (qnumber*)aligned_alloc(16, ((size_t)d * sizeof(qnumber) + 15) / 16 * 16);
	memcpy(out->qsite, qsite, (size_t)d * sizeof(qnumber));
	out->a = (block_sparse_tensor*)calloc((size_t)nsites, sizeof(block_sparse_tensor));
	const qnumber qzero[1] = { 0 };
	for (int l = 0; l < nsites; l++) {
		const int nml = (l == 0 ? 1 : nb), nmr = (l == nsites - 1 ? 1 : nb);
		const int mr0 = (l == nsites - 1 ? nb - 1 : 0);
		const ct_long dim[4] = { nml, d, d, nmr };
		const tensor_axis_direction ad[4] = { TENSOR_AXIS_OUT, TENSOR_AXIS_OUT, TENSOR_AXIS_IN, TENSOR_AXIS_IN };
		const qnumber* qn[4] = { l == 0 ? qzero : qbond, qsite, qsite, l == nsites - 1 ? qzero : qbond };
		block_sparse_tensor* t = &out->a[l];
		allocate_block_sparse_tensor(CT_DOUBLE_REAL, 4, dim, ad, qn, t);
		// scatter the dense entries into the blocks
		const Meta m = meta_of(t);
		std::vector<double> packed((size_t)m.nelem, 0.0);
		ct_long idx[4], bd[4];
		for (ct_long k = 0; k < m.nblocks; k++) {
			if (!m.present(k)) continue;
			m.block_index(k, idx);
			m.block_dims(k, bd);
			std::vector<ct_long> f[4];
			for (int i = 0; i < 4; i++) f[i] = fanout(m, i, m.qnums_blocks[i][idx[i]]);
			ct_long o = m.offset[k];
			for (ct_long a = 0; a < bd[0]; a++) for (ct_long p = 0; p < bd[1]; p++) for (ct_long q = 0; q < bd[2]; q++) for (ct_long b = 0; b < bd[3]; b++)
				packed[(size_t)o++] = w[(size_t)((((f[0][a]) * d + f[1][p]) * d + f[2][q]) * nb + (mr0 + f[3][b]))];
		}
		block_sparse_tensor_deserialize_entries(t, packed.data());
	}
}

static void fsm_put(std::vector<double>& w, int nb, ct_long d, int a, int b, double c, const double* op)
{
	for (ct_long p = 0; p < d; p++) for (ct_long q = 0; q < d; q++) w[(size_t)((((ct_long)a * d + p) * d + q) * nb + b)] += c * op[p * d + q];
}

static qnumber encode_pair(qnumber a, qnumber b) { return a * 65536 + b; }

static void kron2(const double* a, const double* b, double* r)
{
	for (int i1 = 0; i1 < 2; i1++) for (int i2 = 0; i2 < 2; i2++) for (int j1 = 0; j1 < 2; j1++) for (int j2 = 0; j2 < 2; j2++)
		r[(i1 * 2 + i2) * 4 + (j1 * 2 + j2)] = a[i1 * 2 + j1] * b[i2 * 2 + j2];
}

}  // namespace ctg

extern "C" {

void ct_construct_heisenberg_xxz_1d_mpo(const int nsites, const double J, const double D, const double h, struct mpo* mpo)
{
	const ct_long d = 2;
	const qnumber qsite[2] = { 1, -1 };
	const double id[4] = { 1, 0, 0, 1 }, sup[4] = { 0, 1, 0, 0 }, sdn[4] = { 0, 0, 1, 0 }, sz[4] = { 0.5, 0, 0, -0.5 };
	const int nb = 5;
	// states: 0 start, 1 after S+, 2 after S-, 3 after Sz, 4 final
	const qnumber qbond[5] = { 0, 2, -2, 0, 0 };
	std::vector<double> w((size_t)(nb * nb * d * d), 0.0);
	fsm_put(w, nb, d, 0, 0, 1, id);
	fsm_put(w, nb, d, 0, 1, 1, sup);
	fsm_put(w, nb, d, 1, 4, 0.5 * J, sdn);
	fsm_put(w, nb, d, 0, 2, 1, sdn);
	fsm_put(w, nb, d, 2, 4, 0.5 * J, sup);
	fsm_put(w, nb, d, 0, 3, 1, sz);
	fsm_put(w, nb, d, 3, 4, J * D, sz);
	fsm_put(w, nb, d, 0, 4, -h, sz);
	fsm_put(w, nb, d, 4, 4, 1, id);
	mpo_from_fsm(nsites, d, qsite, nb, qbond, w, mpo);
}

void ct_construct_fermi_hubbard_1d_mpo(const int nsites, const double t, const double u, const double mu, struct mpo* mpo)
{
	const ct_long d = 4;
	const qnumber qsite[4] = { encode_pair(0, 0), encode_pair(1, -1), encode_pair(1, 1), encode_pair(2, 0) };
	const double id[4] = { 1, 0, 0, 1 }, cr[4] = { 0, 0, 1, 0 }, an[4] = { 0, 1, 0, 0 }, z[4] = { 1, 0, 0, -1 };
	double Id[16], CZ[16], AZ[16], IC[16], IA[16], CI[16], AI[16], ZA[16], ZC[16], loc[16];
	kron2(id, id, Id); kron2(cr, z, CZ); kron2(an, z, AZ); kron2(id, cr, IC); kron2(id, an, IA);
	kron2(cr, id, CI); kron2(an, id, AI); kron2(z, an, ZA); kron2(z, cr, ZC);
	const double ntot[4] = { 0, 1, 1, 2 }, nint[4] = { 0.25, -0.25, -0.25, 0.25 };
	for (int i = 0; i < 16; i++) loc[i] = 0;
	for (int i = 0; i < 4; i++) loc[i * 4 + i] = -mu * ntot[i] + u * nint[i];
	const int nb = 6;
	// states: 0 start, 1..4 after an open hop (c+_up Z, c_up Z, c+_dn, c_dn), 5 final
	const qnumber qbond[6] = { 0, encode_pair(1, 1), encode_pair(-1, -1), encode_pair(1, -1), encode_pair(-1, 1), 0 };
	std::vector<double> w((size_t)(nb * nb * d * d), 0.0);
	fsm_put(w, nb, d, 0, 0, 1, Id);
	fsm_put(w, nb, d, 0, 1, 1, CZ); fsm_put(w, nb, d, 1, 5, -t, AI);
	fsm_put(w, nb, d, 0, 2, 1, AZ); fsm_put(w, nb, d, 2, 5, -t, CI);
	fsm_put(w, nb, d, 0, 3, 1, IC); fsm_put(w, nb, d, 3, 5, -t, ZA);
	fsm_put(w, nb, d, 0, 4, 1, IA); fsm_put(w, nb, d, 4, 5, -t, ZC);
	fsm_put(w, nb, d, 0, 5, 1, loc);
	fsm_put(w, nb, d, 5, 5, 1, Id);
	mpo_from_fsm(nsites, d, qsite, nb, qbond, w, mpo);
}

/// bond quantum numbers as drawn by construct_random_mps (mps.c:93-150): concatenated lists of
/// the nsites+1 bonds into qbonds_out, their lengths into dims_out
void ct_random_bond_qnums(const int nsites, const ct_long d, const qnumber* qsite, const qnumber sector, const ct_long max_vdim, const uint64_t seed, ct_long* dims_out, qnumber* qbonds_out)
{
	rng_state rng;
	seed_rng_state(seed, &rng);
	std::vector<std::vector<qnumber>> qb;
	random_bond_qnums(nsites, d, qsite, sector, max_vdim, &rng, qb);
	ct_long off = 0;
	for (int l = 0; l <= nsites; l++) {
		dims_out[l] = (ct_long)qb[l].size();
		memcpy(qbonds_out + off, qb[l].data(), qb[l].size() * sizeof(qnumber));
		off += (ct_long)qb[l].size();
	}
}

}  // extern "C"
