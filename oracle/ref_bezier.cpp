// TEST INFRASTRUCTURE ONLY.  Our own driver around the reference's Bezier
// visualisation and temporal VTKHDF headers (include/BezierVisualization.hpp,
// include/TemporalVtkHdf.hpp, with solvers/cpu/include/IgaDatabase.hpp),
// built by `make -C oracle ref-bezier` into oracle/_ref/ref_bezier against the
// HDF5 library of this image (/opt/conda).  tests/golden/make_bezier_reference.py
// runs it to write the reference's report and VTKHDF file for a generated case;
// the digests of both are the fixture tests/golden/bezier_reference.json.
//   ref_bezier DATABASE OUTPUT STATE...
// (the same arguments and outputs as tests/bezier_host_check.cpp)
#include "TemporalVtkHdf.hpp"

#include <cstdio>
#include <fstream>

int main(int argc, char** argv)
{
	try {
		if (argc < 3) throw std::runtime_error("usage: ref_bezier DATABASE OUTPUT STATE...");
		iga::Database database(argv[1]);
		const std::filesystem::path output = argv[2];
		const auto mesh = iga::BuildSourceCoordinateBezierVisualizationMesh(database, false);
		iga::WriteBezierGeometryReport(iga::BezierGeometryReportPath(output), mesh.validation);
		iga::TemporalVtkHdfWriter writer(iga::VtkHdfPath(output), mesh, false);
		const auto nodes = database.header().nodes;
		for (int s = 3; s < argc; ++s) {
			std::ifstream in(argv[s]);
			std::vector<double> v(3*nodes), p(nodes);
			for (std::uint64_t n = 0; n < nodes; ++n)
				if (!(in >> v[3*n] >> v[3*n + 1] >> v[3*n + 2] >> p[n])) throw std::runtime_error("short state file");
			writer.Append(0.001*(s - 2), {{"velocity", 3, v}, {"pressure", 1, p}});
		}
		writer.Close();
		std::printf("points=%zu local=%llu\n", mesh.points.size(),
			static_cast<unsigned long long>(mesh.validation.local_points));
		return 0;
	} catch (const std::exception& e) {
		std::fprintf(stderr, "ref_bezier: %s\n", e.what());
		return 1;
	}
}
