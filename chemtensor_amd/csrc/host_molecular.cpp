// MPO of the spin molecular Hamiltonian (the operator of construct_spin_molecular_hamiltonian_mpo_assembly,
// src/operator/hamiltonian.c:2755-2990), physicists' convention:
//
//   H = sum_{i,j,s} t_ij a+_{i s} a_{j s} + 1/2 sum_{i,j,k,l,s,u} v_ijkl a+_{i s} a+_{j u} a_{l u} a_{k s}
//
// on spatial orbitals with the local basis {0, dn, up, updn} (quantum numbers (N, 2 Sz) = (0,0),
// (1,-1), (1,1), (2,0)). Built here as a finite-state machine over the terms (the reference builds an
// operator graph; the virtual bond basis differs, the operator is identical):
//   - fermionic modes in the Jordan-Wigner order (0 up, 0 dn, 1 up, 1 dn, ...); the operators of a
//     term are ordered by site (the sign of the permutation goes into the coefficient), then the
//     local operator at site k is the product of the term's operators there times P_k to the power
//     of the number of the term's operators right of k (P = local fermion parity);
//   - the state of a term at the bond after site b: nothing placed yet (I), the whole term placed
//     (H), or, on the bonds of the left half, its left operators (one or two, coefficient 1) or,
//     with three placed, its remaining right operator (the coefficient included); on the bonds of
//     the right half its single left operator (coefficient 1) or its remaining right operators (one
//     or two, the coefficient included). Pairs of operators thus index states only on the shorter
//     side of a bond. Terms with the same key share the state; the transition of a term from a state
//     without its coefficient (I, left operators) to one with it (right operators, H) places the
//     coefficient and adds up over the terms, every other transition is fixed by its two states
//     and set once.
// The bond quantum number of a state is the charge of its left part.

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <vector>

#include "ct_internal.h"

namespace ctg {

namespace {

// mode m = 2 * site + spin (spin 0 up, 1 dn); operator code = 2 * mode + dagger
struct Op { int site, spin, dag; };

inline Op op_of(int code) { return Op{ code >> 2, (code >> 1) & 1, code & 1 }; }

// 4 x 4 local matrices (row = output state), basis index 2 * n_up + n_dn
struct Mat4
{
	double a[16];
	Mat4() { std::fill(a, a + 16, 0.0); }
	static Mat4 id() { Mat4 m; for (int i = 0; i < 4; i++) m.a[i * 5] = 1; return m; }
	Mat4 operator*(const Mat4& o) const
	{
		Mat4 r;
		for (int i = 0; i < 4; i++) for (int k = 0; k < 4; k++) {
			const double x = a[i * 4 + k];
			if (x == 0) continue;
			for (int j = 0; j < 4; j++) r.a[i * 4 + j] += x * o.a[k * 4 + j];
		}
		return r;
	}
};

Mat4 kron(const double* x, const double* y)
{
	Mat4 r;
	for (int i1 = 0; i1 < 2; i1++) for (int i2 = 0; i2 < 2; i2++) for (int j1 = 0; j1 < 2; j1++) for (int j2 = 0; j2 < 2; j2++)
		r.a[(i1 * 2 + i2) * 4 + (j1 * 2 + j2)] = x[i1 * 2 + j1] * y[i2 * 2 + j2];
	return r;
}

struct LocalOps
{
	Mat4 c[2][2];   // [spin][dagger]
	Mat4 parity;
	LocalOps()
	{
		const double id[4] = { 1, 0, 0, 1 }, cr[4] = { 0, 0, 1, 0 }, an[4] = { 0, 1, 0, 0 }, z[4] = { 1, 0, 0, -1 };
		c[0][0] = kron(an, id);   // up: first mode of the site
		c[0][1] = kron(cr, id);
		c[1][0] = kron(z, an);    // dn: behind the up mode of the same site
		c[1][1] = kron(z, cr);
		parity = kron(z, z);
	}
};

// charge (N, 2 Sz) of an operator
inline void op_charge(const Op& o, int& dn, int& ds)
{
	const int s = (o.spin == 0) ? 1 : -1;
	dn = o.dag ? 1 : -1;
	ds = o.dag ? s : -s;
}

// state keys: [kind, ops...]; kind 0 = I, 1 = left operators (their codes in order), 2 = three
// operators applied (the code of the remaining one), 3 = H
using Key = std::vector<int>;

}  // namespace

}  // namespace ctg

extern "C" void ct_construct_spin_molecular_hamiltonian_mpo(const int nsites, const double* tkin, const double* vint, struct mpo* mpo)
{
	using namespace ctg;
	const int n = nsites;
	const LocalOps L;
	// bond b = state after site b (b = -1 .. n-1); per bond the state index of each key
	std::vector<std::map<Key, int>> states((size_t)n + 1);
	std::vector<std::vector<std::pair<int, int>>> charge((size_t)n + 1);   // per bond and state: (N, 2 Sz)
	auto state = [&](int b, const Key& k, int qn, int qs) {
		std::map<Key, int>& m = states[(size_t)(b + 1)];
		auto it = m.find(k);
		if (it != m.end()) return it->second;
		const int id = (int)m.size();
		m.emplace(k, id);
		charge[(size_t)(b + 1)].push_back({ qn, qs });
		return id;
	};
	// transitions at site k: (left state, right state) -> local matrix
	std::vector<std::map<std::pair<int, int>, Mat4>> trans((size_t)n);
	std::vector<std::map<std::pair<int, int>, bool>> pass((size_t)n);   // pass-through transitions (set once)

	const Key kI = { 0 }, kH = { 3 };
	auto add_term = [&](double coeff, std::vector<int> ops) {
		if (coeff == 0) return;
		// a product with the same creation (annihilation) operator twice vanishes
		for (size_t x = 0; x + 1 < ops.size(); x++)
			for (size_t y = x + 1; y < ops.size(); y++)
				if (ops[x] == ops[y]) {
					bool between = false;   // a+_x a_x a+_x ... only vanishes if nothing of the same mode sits between
					for (size_t z = x + 1; z < y; z++) if ((ops[z] >> 1) == (ops[x] >> 1)) between = true;
					if (!between) return;
				}
		// stable order by site; the sign of the permutation into the coefficient
		const int m = (int)ops.size();
		for (int x = 1; x < m; x++)
			for (int y = x; y > 0 && (ops[y] >> 2) < (ops[y - 1] >> 2); y--) {
				std::swap(ops[y], ops[y - 1]);
				coeff = -coeff;
			}
		// two creators (or two annihilators) at one site: spin up first, so that a+_up a+_dn and
		// a+_dn a+_up (distinct modes, anticommuting) share one state, as the reference's graph
		// keys its pair chains by ordered (i, s) < (j, u)
		for (int x = 0; x + 1 < m; x++) {
			const Op p = op_of(ops[x]), q = op_of(ops[x + 1]);
			if (p.site == q.site && p.dag == q.dag && p.spin > q.spin) {
				std::swap(ops[x], ops[x + 1]);
				coeff = -coeff;
			}
		}
		int cn = 0, cs = 0;   // charge of the left part
		// the state of the term after site b with 'cnt' of its operators left of the bond
		auto key_at = [&](int bond, int cnt) -> Key {
			if (cnt == 0) return kI;
			if (cnt == m) return kH;
			const bool left_keyed = (bond + 1) <= (n - 1 - bond);
			Key k;
			if (m == 2) {
				if (left_keyed) k = Key{ 1, ops[0] };
				else k = Key{ 2, ops[1] };
			}
			else if (left_keyed) {
				if (cnt <= 2) { k = Key{ 1 }; k.insert(k.end(), ops.begin(), ops.begin() + cnt); }
				else k = Key{ 2, ops[3] };
			}
			else {
				if (cnt == 1) k = Key{ 1, ops[0] };
				else { k = Key{ 2 }; k.insert(k.end(), ops.begin() + cnt, ops.end()); }
			}
			return k;
		};
		// kinds 0 (I) and 1 (left operators) carry no coefficient, 2 (right operators) and 3 (H) do
		auto weighted_kind = [](const Key& k) { return k[0] >= 2; };
		Key pk = kI;
		int prev = state(-1, kI, 0, 0);
		int count = 0;
		int x = 0;
		for (int k = 0; k < n; k++) {
			const int x0 = x;
			while (x < m && (ops[x] >> 2) == k) x++;
			const int mk = x - x0;
			Mat4 loc = Mat4::id();
			for (int y = x0; y < x; y++) {
				const Op o = op_of(ops[y]);
				loc = loc * L.c[o.spin][o.dag];
				int dn, ds;
				op_charge(o, dn, ds);
				cn += dn;
				cs += ds;
			}
			count += mk;
			if (count % 2) loc = loc * L.parity;
			const Key nk = key_at(k, count);
			const int next = state(k, nk, cn, cs);
			const auto key = std::make_pair(prev, next);
			if (!weighted_kind(pk) && weighted_kind(nk)) {
				// the one transition of the term that places its coefficient: the terms add up
				Mat4& t = trans[(size_t)k][key];
				for (int e = 0; e < 16; e++) t.a[e] += coeff * loc.a[e];
			}
			else if (!pass[(size_t)k].count(key)) {
				// every other transition is fixed by its two states: set once
				pass[(size_t)k][key] = true;
				trans[(size_t)k][key] = loc;
			}
			pk = nk;
			prev = next;
		}
	};

	// I -> I and H -> H at every site (the only states of the outer bonds)
	for (int k = 0; k < n; k++) {
		const int a = state(k - 1, kI, 0, 0), b = state(k, kI, 0, 0);
		trans[(size_t)k][std::make_pair(a, b)] = Mat4::id();
		pass[(size_t)k][std::make_pair(a, b)] = true;
		if (k > 0) {
			const int h0 = state(k - 1, kH, 0, 0), h1 = state(k, kH, 0, 0);
			trans[(size_t)k][std::make_pair(h0, h1)] = Mat4::id();
			pass[(size_t)k][std::make_pair(h0, h1)] = true;
		}
		else state(k, kH, 0, 0);
	}
	for (int i = 0; i < n; i++)
		for (int j = 0; j < n; j++)
			for (int s = 0; s < 2; s++)
				add_term(tkin[i * n + j], { (((i * 2) + s) * 2) + 1, (((j * 2) + s) * 2) + 0 });
	for (int i = 0; i < n; i++) for (int j = 0; j < n; j++) for (int k = 0; k < n; k++) for (int l = 0; l < n; l++) {
		const double v = vint[((i * n + j) * n + k) * n + l];
		if (v == 0) continue;
		for (int s = 0; s < 2; s++) for (int u = 0; u < 2; u++)
			add_term(0.5 * v, { ((i * 2 + s) * 2) + 1, ((j * 2 + u) * 2) + 1, ((l * 2 + u) * 2) + 0, ((k * 2 + s) * 2) + 0 });
	}

	// assemble: site k maps the states of bond k-1 (rows) to those of bond k (columns); the outer bonds
	// keep only I (left) and H (right)
	mpo->nsites = n;
	mpo->d = 4;
	mpo->qsite = (qnumber*)aligned_alloc(16, 16);
	const qnumber qsite[4] = { 0, 1 * 65536 - 1, 1 * 65536 + 1, 2 * 65536 };
	memcpy(mpo->qsite, qsite, sizeof(qsite));
	mpo->a = (block_sparse_tensor*)calloc((size_t)n, sizeof(block_sparse_tensor));
	auto bond_list = [&](int b, std::vector<int>& keep) {
		// b = -1: only I; b = n-1: only H; otherwise every state
		keep.clear();
		const std::map<Key, int>& m = states[(size_t)(b + 1)];
		if (b == -1) { keep.push_back(m.at(kI)); return; }
		if (b == n - 1) { keep.push_back(m.at(kH)); return; }
		for (int s = 0; s < (int)m.size(); s++) keep.push_back(s);
	};
	for (int k = 0; k < n; k++) {
		std::vector<int> lk, rk;
		bond_list(k - 1, lk);
		bond_list(k, rk);
		std::vector<int> lpos(states[(size_t)k].size(), -1), rpos(states[(size_t)(k + 1)].size(), -1);
		for (size_t x = 0; x < lk.size(); x++) lpos[(size_t)lk[x]] = (int)x;
		for (size_t x = 0; x < rk.size(); x++) rpos[(size_t)rk[x]] = (int)x;
		std::vector<qnumber> ql(lk.size()), qr(rk.size());
		for (size_t x = 0; x < lk.size(); x++) { const auto& c = charge[(size_t)k][(size_t)lk[x]]; ql[x] = c.first * 65536 + c.second; }
		for (size_t x = 0; x < rk.size(); x++) { const auto& c = charge[(size_t)(k + 1)][(size_t)rk[x]]; qr[x] = c.first * 65536 + c.second; }
		const ct_long dim[4] = { (ct_long)lk.size(), 4, 4, (ct_long)rk.size() };
		const tensor_axis_direction ad[4] = { TENSOR_AXIS_OUT, TENSOR_AXIS_OUT, TENSOR_AXIS_IN, TENSOR_AXIS_IN };
		const qnumber* qn[4] = { ql.data(), qsite, qsite, qr.data() };
		block_sparse_tensor* t = &mpo->a[k];
		allocate_block_sparse_tensor(CT_DOUBLE_REAL, 4, dim, ad, qn, t);
		const Meta mt = meta_of(t);
		// logical index -> (block number on the axis, position within the block)
		std::vector<std::vector<ct_long>> bno(4), bpos(4);
		for (int ax = 0; ax < 4; ax++) {
			bno[ax].resize((size_t)mt.dim_logical[ax]);
			bpos[ax].resize((size_t)mt.dim_logical[ax]);
			std::vector<ct_long> cnt((size_t)mt.dim_blocks[ax], 0);
			for (ct_long x = 0; x < mt.dim_logical[ax]; x++) {
				const ct_long b = mt.find_qnum(ax, mt.qnums_logical[ax][x]);
				bno[ax][x] = b;
				bpos[ax][x] = cnt[b]++;
			}
		}
		std::vector<double> packed((size_t)mt.nelem, 0.0);
		for (const auto& tr : trans[(size_t)k]) {
			const int a = lpos[(size_t)tr.first.first], b = rpos[(size_t)tr.first.second];
			if (a < 0 || b < 0) continue;
			for (int p = 0; p < 4; p++) for (int q = 0; q < 4; q++) {
				const double v = tr.second.a[p * 4 + q];
				if (v == 0) continue;
				const ct_long ib[4] = { bno[0][a], bno[1][p], bno[2][q], bno[3][b] };
				const ct_long kb = mt.flat_index(ib);
				if (!mt.present(kb)) continue;   // forbidden by the charges: the entry is zero
				ct_long bd[4];
				mt.block_dims(kb, bd);
				const ct_long off = mt.offset[kb] + ((bpos[0][a] * bd[1] + bpos[1][p]) * bd[2] + bpos[2][q]) * bd[3] + bpos[3][b];
				packed[(size_t)off] += v;
			}
		}
		block_sparse_tensor_deserialize_entries(t, packed.data());
	}
}
