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
		tfi_host::OutletKind kind;
		if (!tfi_host::ParseOutletKind(argv[1], &kind)) throw std::runtime_error("unknown outlet kind");
		const double dt = std::stod(argv[9]), flow = std::stod(argv[10]);
		tfi_host::OutletBank bank;
		bank.Add(2, kind, std::stod(argv[2]), std::stod(argv[3]), std::stod(argv[4]), std::stod(argv[5]),
			std::stod(argv[6]), std::stod(argv[7]), std::stod(argv[8]));
		tfi_host::OutletBank relaxed = bank, committed = bank;
		bank.StartStep();
		const double change = bank.Respond({flow}, dt);
		relaxed.StartStep();
		relaxed.Respond({flow}, dt);
		relaxed.Relax();
		committed.StartStep();
		committed.Respond({flow}, dt);
		committed.Commit();
		// the model's own response: the committed pressure and capacitor pressure
		for (double v : {committed.applied(0), committed.capacitor(0), change, bank.Tolerance(), relaxed.applied(0),
			committed.applied(0), committed.capacitor(0)})
			std::printf("%.17g\n", v);
		return 0;
	} catch (const std::exception& ex) {
		std::fprintf(stderr, "%s\n", ex.what());
		return 1;
	}
}
