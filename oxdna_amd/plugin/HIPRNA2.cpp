/*
 * HIPRNA2: an interaction plugin for the reference oxDNA program that
 * evaluates the oxRNA2 forces on an AMD GPU through liboxh (include/oxh.h),
 * as HIPDNA2 does for oxDNA2 (oxdna_amd/plugin/HIPDNA2.cpp, which describes
 * how the forces reach the reference's MD_CPUBackend).
 *
 *     interaction_type = HIPRNA2
 *     plugin_search_path = <folder of HIPRNA2.so>
 *     backend_precision = mixed | double     (GPU outputs; the oxRNA terms are FP64)
 *     CUDA_device = 0
 *
 * The class is the reference's RNA2Interaction (src/Interactions/
 * RNAInteraction2.cpp) with the force evaluation of a whole step moved to
 * the GPU: the system's state, box, temperature, salt and Debye-Hueckel
 * settings and max_backbone_force go to oxh_system_desc with interaction =
 * OXH_INTERACTION_RNA2.  The GPU evaluates the hydrogen-bonding f4 factors
 * analytically, as the reference's GPU kernel does; the CPU pair calls
 * made for observables (update_forces false) keep the reference's meshed
 * ones.  Refused: sequence-dependent parameters, mismatch_repulsion, and
 * systems above 1,024 nucleotides (the GPU's oxRNA path takes systems
 * without cell lists).
 */

#include <dlfcn.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include <hip/hip_runtime_api.h>

#include "Boxes/BaseBox.h"
#include "Interactions/RNAInteraction2.h"
#include "Particles/BaseParticle.h"
#include "Utilities/ConfigInfo.h"
#include "Utilities/Utils.h"
#include "Utilities/oxDNAException.h"

#include "oxh.h"

namespace {

// The HIP runtime's start-up looks for an optional profiler hook with
// dlsym and leaves that lookup's "undefined symbol" message in the dlerror
// state.  The reference's PluginManager reads dlerror() right after dlopen
// (src/PluginManagement/PluginManager.cpp:111-115) and would take the
// message for a failed load.  A library's constructors run after those of
// the libraries it needs, so this one clears the message.
__attribute__((constructor)) void clear_dlerror_of_dependencies() {
	(void) dlerror();
}

// (x, y, z, w) unit quaternion of the rotation whose matrix has the columns
// a1, a2, a3 (Shepperd's method: the largest of 1 + trace and 1 + 2 M_kk -
// trace picks the component taken from a square root)
void quat_of(const LR_vector &a1, const LR_vector &a2, const LR_vector &a3, double *q) {
	const double M[3][3] = {{a1.x, a2.x, a3.x}, {a1.y, a2.y, a3.y}, {a1.z, a2.z, a3.z}};
	const double tr = M[0][0] + M[1][1] + M[2][2];
	int k = -1;
	double best = tr;
	for(int i = 0; i < 3; i++)
		if(M[i][i] > best) {
			best = M[i][i];
			k = i;
		}
	if(k < 0) {
		const double s = 2.0 * std::sqrt(1.0 + tr);
		q[0] = (M[2][1] - M[1][2]) / s;
		q[1] = (M[0][2] - M[2][0]) / s;
		q[2] = (M[1][0] - M[0][1]) / s;
		q[3] = 0.25 * s;
	}
	else {
		const int j = (k + 1) % 3, l = (k + 2) % 3;
		const double s = 2.0 * std::sqrt(1.0 + M[k][k] - M[j][j] - M[l][l]);
		q[k] = 0.25 * s;
		q[j] = (M[j][k] + M[k][j]) / s;
		q[l] = (M[l][k] + M[k][l]) / s;
		q[3] = (M[l][j] - M[j][l]) / s;
	}
	const double n = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
	for(int c = 0; c < 4; c++) q[c] /= n;
}

}  // namespace

class HIPRNA2Interaction: public RNA2Interaction {
protected:
	oxh_batch *_batch = nullptr;
	int _precision = 0;   // 0 mixed, 1 double
	int _device = 0;
	int _N = 0;
	double _box[3] = {0., 0., 0.};
	double *_d_pos = nullptr, *_d_quat = nullptr;
	void *_d_F = nullptr, *_d_T = nullptr;
	std::vector<double> _h_pos, _h_quat, _gF, _gT;
	std::vector<float> _Ff, _Tf;
	std::vector<char> _pending;
	bool _any_pending = false;

	void _check(oxh_status s, const char *what) {
		if(s != OXH_OK) throw oxDNAException("HIPRNA2: %s failed: %s", what, oxh_last_error());
	}
	void _hip(hipError_t e, const char *what) {
		if(e != hipSuccess) throw oxDNAException("HIPRNA2: %s failed: %s", what, hipGetErrorString(e));
	}

	// the GPU copy of the system, from the particles and this interaction's
	// parameters (read by get_settings / init as for RNA2Interaction)
	void _create(std::vector<BaseParticle *> &particles) {
		_N = (int) particles.size();
		std::vector<double> pos(3 * _N), a1(3 * _N), a3(3 * _N);
		std::vector<int> btype(_N), n3(_N), n5(_N);
		for(auto p : particles) {
			const int i = p->index;
			const LR_vector &v1 = p->orientationT.v1, &v3 = p->orientationT.v3;
			const double x[3] = {p->pos.x, p->pos.y, p->pos.z}, u[3] = {v1.x, v1.y, v1.z}, w[3] = {v3.x, v3.y, v3.z};
			for(int c = 0; c < 3; c++) {
				pos[3 * i + c] = x[c];
				a1[3 * i + c] = u[c];
				a3[3 * i + c] = w[c];
			}
			btype[i] = p->btype;
			n3[i] = p->n3 != P_VIRTUAL ? p->n3->index : -1;
			n5[i] = p->n5 != P_VIRTUAL ? p->n5->index : -1;
		}
		oxh_system_desc d;
		oxh_system_desc_defaults(&d);
		d.N = _N;
		d.pos = pos.data();
		d.a1 = a1.data();
		d.a3 = a3.data();
		d.btype = btype.data();
		d.n3 = n3.data();
		d.n5 = n5.data();
		LR_vector sides = CONFIG_INFO->box->box_sides();
		for(int c = 0; c < 3; c++) d.box[c] = _box[c] = sides[c];
		d.T = _T;
		d.salt = _salt_concentration;
		d.dh_half_charged_ends = _debye_huckel_half_charged_ends ? 1 : 0;
		d.dh_lambda = _debye_huckel_lambdafactor;
		d.dh_strength = _debye_huckel_prefactor;
		d.dh_rhigh = _debye_huckel_RHIGH;
		if(_use_mbf) {
			d.max_backbone_force = _mbf_fmax;
			d.max_backbone_force_far = _mbf_finf;
		}
		d.backend_precision = _precision;
		d.interaction = OXH_INTERACTION_RNA2;
		_check(oxh_batch_create(_device, 1, &d, &_batch), "oxh_batch_create");
		const size_t nb = (size_t) 4 * _N * (_precision ? sizeof(double) : sizeof(float));
		_hip(hipMalloc((void **) &_d_pos, (size_t) 4 * _N * sizeof(double)), "hipMalloc");
		_hip(hipMalloc((void **) &_d_quat, (size_t) 4 * _N * sizeof(double)), "hipMalloc");
		_hip(hipMalloc(&_d_F, nb), "hipMalloc");
		_hip(hipMalloc(&_d_T, nb), "hipMalloc");
		_h_pos.assign(4 * _N, 0.0);
		_h_quat.assign(4 * _N, 0.0);
		_gF.assign(4 * _N, 0.0);
		_gT.assign(4 * _N, 0.0);
		_pending.assign(_N, 0);
	}

	// forces, torques and energy shares of the current configuration into
	// _gF / _gT (4 per particle, upload order = particle index)
	void _gpu_forces() {
		std::vector<BaseParticle *> &particles = CONFIG_INFO->particles();
		if(!_batch) _create(particles);
		if((int) particles.size() != _N) throw oxDNAException("HIPRNA2: the number of particles changed");
		LR_vector sides = CONFIG_INFO->box->box_sides();
		for(int c = 0; c < 3; c++)
			if(sides[c] != _box[c]) throw oxDNAException("HIPRNA2: the box changed; the GPU copy keeps the box of the first step");
		for(auto p : particles) {
			const int i = p->index;
			for(int c = 0; c < 3; c++) _h_pos[4 * i + c] = p->pos[c];
			quat_of(p->orientationT.v1, p->orientationT.v2, p->orientationT.v3, &_h_quat[4 * i]);
		}
		const size_t nb = (size_t) 4 * _N * (_precision ? sizeof(double) : sizeof(float));
		_hip(hipMemcpy(_d_pos, _h_pos.data(), (size_t) 4 * _N * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy");
		_hip(hipMemcpy(_d_quat, _h_quat.data(), (size_t) 4 * _N * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy");
		_hip(hipMemset(_d_F, 0, nb), "hipMemset");
		_hip(hipMemset(_d_T, 0, nb), "hipMemset");
		_check(oxh_interaction_compute_forces(_batch, _d_pos, _d_quat, _d_F, _d_T, nullptr), "oxh_interaction_compute_forces");
		_check(oxh_batch_sync(_batch), "oxh_batch_sync");
		if(_precision) {
			_hip(hipMemcpy(_gF.data(), _d_F, nb, hipMemcpyDeviceToHost), "hipMemcpy");
			_hip(hipMemcpy(_gT.data(), _d_T, nb, hipMemcpyDeviceToHost), "hipMemcpy");
		}
		else {
			_Ff.resize(4 * _N);
			_Tf.resize(4 * _N);
			_hip(hipMemcpy(_Ff.data(), _d_F, nb, hipMemcpyDeviceToHost), "hipMemcpy");
			_hip(hipMemcpy(_Tf.data(), _d_T, nb, hipMemcpyDeviceToHost), "hipMemcpy");
			for(int k = 0; k < 4 * _N; k++) {
				_gF[k] = _Ff[k];
				_gT[k] = _Tf[k];
			}
		}
	}

	// particle p's GPU force, torque and energy share, the first time p
	// appears in the backend's pair loop
	number _hand_over(BaseParticle *p) {
		const int i = p->index;
		if(!_pending[i]) return (number) 0.;
		_pending[i] = 0;
		p->force += LR_vector(_gF[4 * i], _gF[4 * i + 1], _gF[4 * i + 2]);
		p->torque += LR_vector(_gT[4 * i], _gT[4 * i + 1], _gT[4 * i + 2]);
		return (number) _gF[4 * i + 3];
	}

public:
	HIPRNA2Interaction() = default;

	virtual ~HIPRNA2Interaction() {
		if(_batch) oxh_batch_destroy(_batch);
		if(_d_pos) (void) hipFree(_d_pos);
		if(_d_quat) (void) hipFree(_d_quat);
		if(_d_F) (void) hipFree(_d_F);
		if(_d_T) (void) hipFree(_d_T);
	}

	void get_settings(input_file &inp) override {
		RNA2Interaction::get_settings(inp);
		std::string prec("mixed");
		getInputString(&inp, "backend_precision", prec, 0);
		if(prec == "mixed") _precision = 0;
		else if(prec == "double") _precision = 1;
		else throw oxDNAException("HIPRNA2: backend_precision = %s; the GPU path takes mixed or double", prec.c_str());
		getInputInt(&inp, "CUDA_device", &_device, 0);
		if(!_average) throw oxDNAException("HIPRNA2: the GPU path takes the average-sequence oxRNA parameters only");
		if(_mismatch_repulsion) throw oxDNAException("HIPRNA2: mismatch_repulsion is not available on the GPU path");
	}

	void begin_energy_and_force_computation() override {
		RNA2Interaction::begin_energy_and_force_computation();
		_gpu_forces();
		std::fill(_pending.begin(), _pending.end(), 1);
		_any_pending = true;
	}

	number pair_interaction_bonded(BaseParticle *p, BaseParticle *q, bool compute_r = true, bool update_forces = false) override {
		if(update_forces && _any_pending) return _hand_over(p) + _hand_over(q);
		return RNA2Interaction::pair_interaction_bonded(p, q, compute_r, update_forces);
	}

	number pair_interaction_nonbonded(BaseParticle *p, BaseParticle *q, bool compute_r = true, bool update_forces = false) override {
		if(update_forces && _any_pending) return _hand_over(p) + _hand_over(q);
		return RNA2Interaction::pair_interaction_nonbonded(p, q, compute_r, update_forces);
	}

	number get_system_energy(std::vector<BaseParticle *> &particles, BaseList *lists) override {
		_any_pending = false;
		_gpu_forces();
		double U = 0.;
		for(int i = 0; i < _N; i++) U += _gF[4 * i + 3];
		return (number) U;
	}

	std::map<int, number> get_system_energy_split(std::vector<BaseParticle *> &particles, BaseList *lists) override {
		_any_pending = false;
		_gpu_forces();
		double E8[8];
		_check(oxh_batch_energy_split(_batch, 0, E8), "oxh_batch_energy_split");
		std::map<int, number> out;
		// the GPU's order is RNA2Interaction's: FENE, bonded excluded volume,
		// stacking, non-bonded excluded volume, HB, cross and coaxial
		// stacking, Debye-Hueckel
		const int names[8] = {BACKBONE, BONDED_EXCLUDED_VOLUME, STACKING, NONBONDED_EXCLUDED_VOLUME, HYDROGEN_BONDING, CROSS_STACKING,
				COAXIAL_STACKING, DEBYE_HUCKEL};
		for(int k = 0; k < 8; k++) out[names[k]] = (number) E8[k];
		return out;
	}
};

extern "C" BaseInteraction *make_HIPRNA2() {
	return new HIPRNA2Interaction();
}
