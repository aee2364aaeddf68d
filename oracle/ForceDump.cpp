/*
 * ForceDump: an observable plugin for the reference oxDNA CPU binary.
 *
 * Test infrastructure only.  It is compiled against the reference's headers
 * (oracle/ref.mk) into oracle/_ref/ForceDump.so and loaded by the reference's
 * own PluginManager (src/PluginManagement/PluginManager.cpp:161-180, entry
 * point make_<name>) through `plugin_search_path`.  It prints, for every
 * particle, the force and the torque that the reference's MD backend holds,
 * at full precision (%.17g), because the reference's ParticleForceAndTorque
 * observable prints only 6 significant digits
 * (src/Observables/ParticleForceAndTorque.cpp:52).
 *
 * Line format:  index Fx Fy Fz Tx Ty Tz
 * The torque is in the particle's body frame, as the reference stores it
 * (p->torque += p->orientationT * torque_lab, src/Interactions/DNAInteraction.cpp:461-462).
 */

#include "Observables/BaseObservable.h"
#include "Particles/BaseParticle.h"
#include "Utilities/Utils.h"

#include <cstdio>
#include <string>

class ForceDump: public BaseObservable {
public:
	ForceDump() = default;
	~ForceDump() override = default;

	std::string get_output_string(llint curr_step) override {
		std::string out;
		char buf[512];
		for(auto p : _config_info->particles()) {
			snprintf(buf, sizeof(buf), "%d %.17g %.17g %.17g %.17g %.17g %.17g\n", p->index,
					(double) p->force.x, (double) p->force.y, (double) p->force.z,
					(double) p->torque.x, (double) p->torque.y, (double) p->torque.z);
			out += buf;
		}
		return out;
	}
};

extern "C" BaseObservable *make_ForceDump() {
	return new ForceDump();
}
