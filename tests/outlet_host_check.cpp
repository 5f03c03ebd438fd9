// Test driver for the drop-in driver's outlet models (host/outlets.hpp), built
// with g++ by tests/test_oracle_outlets.py.  Same arguments and output as the
// reference-side probe's `outlet` command (oracle/ref_probe.cpp):
//   outlet_host_check KIND R RP RD C PREF PCAP PRESSURE DT FLOW
// prints the model's pressure and capacitor pressure, the coupling's maximum
// pressure change and tolerance, the relaxed pressure and the committed
// pressure and capacitor pressure, 17 significant digits each.
#include "outlets.hpp"

#include <cstdio>
#include <string>

int main(int argc, char** argv)
{
	try {
		if (argc != 11) throw std::runtime_error("usage: outlet_host_check KIND R RP RD C PREF PCAP PRESSURE DT FLOW");
		tfi_host::OutletModel m;
		m.label = 2;
		m.kind = argv[1];
		if (!tfi_host::IsOutletKind(m.kind)) throw std::runtime_error("unknown outlet kind");
		m.resistance = std::stod(argv[2]);
		m.proximal_resistance = std::stod(argv[3]);
		m.distal_resistance = std::stod(argv[4]);
		m.capacitance = std::stod(argv[5]);
		m.reference_pressure = std::stod(argv[6]);
		m.capacitor_pressure = std::stod(argv[7]);
		m.pressure = std::stod(argv[8]);
		const double dt = std::stod(argv[9]), flow = std::stod(argv[10]);
		const auto e = tfi_host::EvaluateOutlet(m, m.capacitor_pressure, flow, dt);
		const auto c = tfi_host::EvaluateOutlets({m}, {m.capacitor_pressure}, {flow}, dt);
		std::vector<tfi_host::OutletModel> relaxed{m}, committed{m};
		tfi_host::RelaxOutlets(relaxed, c);
		tfi_host::CommitOutlets(committed, c);
		for (double v : {e.pressure, e.capacitor_pressure, c.maximum_pressure_change, c.Tolerance(), relaxed[0].pressure,
			committed[0].pressure, committed[0].capacitor_pressure})
			std::printf("%.17g\n", v);
		return 0;
	} catch (const std::exception& ex) {
		std::fprintf(stderr, "%s\n", ex.what());
		return 1;
	}
}
