// Test driver for the drop-in driver's Bezier VTKHDF output
// (tubularflowiga_amd/host/bezier_output.hpp), built with g++ by
// tests/test_bezier_output.py.  Same arguments and outputs as the
// reference-side probe oracle/ref_bezier.cpp:
//   bezier_host_check DATABASE OUTPUT STATE...
// builds the merged Bezier mesh in source coordinates, writes
// OUTPUT.bezier_geometry.json and OUTPUT's .vtkhdf, appends each STATE file
// ("vx vy vz p" per control point) as "velocity" and "pressure" at times
// 0.001, 0.002, ..., and prints the point and reference counts.
//   bezier_host_check --resume FIRST DATABASE OUTPUT STATE...
// opens the existing file instead (a restarted run) and appends the states
// at times 0.001 FIRST, 0.001 (FIRST + 1), ...
#include "bezier_output.hpp"

#include <cstdio>
#include <fstream>
#include <string>

int main(int argc, char** argv)
{
	try {
		bool resume = false;
		int first = 1;
		if (argc > 2 && std::string(argv[1]) == "--resume") {
			resume = true;
			first = std::stoi(argv[2]);
			argv += 2;
			argc -= 2;
		}
		if (argc < 3) throw std::runtime_error("usage: bezier_host_check [--resume FIRST] DATABASE OUTPUT STATE...");
		const auto db = tfi_host::ReadDatabase(argv[1]);
		const std::filesystem::path output = argv[2];
		const auto mesh = tfi_host::BuildBezierVisualMesh(db);
		tfi_host::WriteBezierReport(tfi_host::BezierReportPath(output), mesh.checks);
		tfi_host::BezierVtkHdfSeries series(tfi_host::VtkHdfOutputPath(output), mesh, resume);
		for (int s = 3; s < argc; ++s) {
			std::ifstream in(argv[s]);
			std::vector<double> v(3*db.nodes), p(db.nodes);
			for (std::uint64_t n = 0; n < db.nodes; ++n)
				if (!(in >> v[3*n] >> v[3*n + 1] >> v[3*n + 2] >> p[n])) throw std::runtime_error("short state file");
			series.Append(0.001*(s - 3 + first), {{"velocity", 3, v}, {"pressure", 1, p}});
		}
		series.Close();
		std::printf("points=%zu local=%llu\n", mesh.points.size(), static_cast<unsigned long long>(mesh.checks.local_points));
		return 0;
	} catch (const std::exception& e) {
		std::fprintf(stderr, "bezier_host_check: %s\n", e.what());
		return 1;
	}
}
