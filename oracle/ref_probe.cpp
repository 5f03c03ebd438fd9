// TEST INFRASTRUCTURE ONLY.  Our own driver around the reference's PETSc-free
// headers (solvers/cpu/include/{IgaDatabase,ElementGeometry,Quadrature,
// BoundaryFlow,BoundarySupport}.hpp), built by `make -C oracle ref` into
// oracle/_ref/ref_probe.  It lets tests/golden/make_ref_vectors.py write
// reference-computed vectors that pin the CPU oracle.  Nothing here ships.
//
// Commands (all numbers printed with 17 significant digits):
//   ref_probe basis   ELEMENT u v w      -> det, inverse(9), per basis: value grad(3)
//   ref_probe flow    ELEMENT face       -> IntegrateBoundaryFlow (state from file)
//   ref_probe traction ELEMENT face p    -> IntegrateBoundaryPressureTraction (4*nen)
//   ref_probe surface ELEMENT face       -> SurfaceJacobianMeasure at the 16 face points
//   ref_probe ntiga   DATABASE           -> header + every element record as text
//   ref_probe walltrace DATABASE VTK     -> WallTraceBasis node ids
//   ref_probe outlet-config CONFIG.json  -> InitializeOutletModels of the first
//                                          system: per model label, kind index
//                                          (0 resistance, 1 RC, 2 RCR), pressure,
//                                          capacitor pressure; then the traction
//                                          value MaterializeOutletPressures gives
//                                          each outlet label
//   ref_probe outlet KIND R RP RD C PREF PCAP PRESSURE DT FLOW
//                                        -> EvaluateOutletModel (pressure, capacitor
//                                          pressure); EvaluateOutletCoupling with
//                                          PCAP as the step's start (maximum change,
//                                          tolerance); pressure after
//                                          RelaxOutletCoupling; pressure and capacitor
//                                          pressure after CommitOutletCoupling
//                                          (include/OutletModel.hpp)
//   ref_probe transport-case CASE_DIR NODES -> ReadTransportParameters of
//                                          simulation_parameter.txt, then
//                                          ResolveBoundaryConditions with
//                                          ReadCaseConfiguration (case_config.json,
//                                          optional) and ReadPointLabels: the
//                                          twelve parameters, the transport node
//                                          count, the sum of the constrained node
//                                          indices, the sums of N0 and Nplus and
//                                          the velocity / pressure node counts
//                                          (solvers/cpu/include/CaseInput.hpp)
//   ref_probe solve-case CASE_DIR DATABASE [SYSTEM]
//                                        -> CompileLinearSystem of the named (or
//                                          first) linear_transport system, the
//                                          per-field SUPG code of
//                                          ConfiguredTransportHost.inc:186-193,
//                                          SupgTauDiffusivity, InitialScalarValues,
//                                          and ResolveScalarBoundaries with
//                                          NodeFaceLabels: "term", "field" and
//                                          "boundary" lines as `iga_hip solve-check`;
//                                          with CASE_DIR/physiology_state.txt (an
//                                          interleaved state), ComputePhysiologyPointArrays
//                                          as "derived" lines
// ELEMENT file: nen, then nen connectivity ints, nen x 64 extraction values,
// 64 x 3 Bezier points, nen x 4 nodal state values (whitespace separated).
#include "BoundaryFlow.hpp"
#include "BoundarySupport.hpp"
#include "IgaDatabase.hpp"
#include "Quadrature.hpp"
#include "OutletModel.hpp"
#include "CaseInput.hpp"
#include "GenericCaseInput.hpp"
#include "PhysiologyOutput.hpp"

#include <cstdio>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

namespace {

struct Input {
	iga::Element element;
	std::vector<std::array<double, 4>> state;
};

Input ReadElement(const char* path)
{
	std::ifstream in(path);
	if (!in) throw std::runtime_error("cannot open element file");
	Input r;
	std::size_t nen = 0;
	in >> nen;
	r.element.connectivity.resize(nen);
	r.element.extraction.resize(nen);
	for (auto& c : r.element.connectivity) in >> c;
	for (auto& row : r.element.extraction)
		for (double& v : row) in >> v;
	for (auto& p : r.element.bezier_points)
		for (double& v : p) in >> v;
	r.state.resize(nen);
	for (auto& s : r.state)
		for (double& v : s) in >> v;
	if (!in) throw std::runtime_error("element file is truncated");
	return r;
}

void Print(double v) { std::printf("%.17g\n", v); }

}  // namespace

int main(int argc, char** argv)
{
	try {
		if (argc < 3) throw std::runtime_error("usage: ref_probe COMMAND ...");
		const std::string cmd = argv[1];
		if (cmd == "basis" && argc == 6) {
			const auto in = ReadElement(argv[2]);
			const auto b = iga::EvaluateBoundaryBasis(in.element, std::stod(argv[3]), std::stod(argv[4]),
				std::stod(argv[5]));
			Print(b.raw_determinant);
			for (int i = 0; i < 3; ++i)
				for (int j = 0; j < 3; ++j) Print(b.inverse_jacobian[i][j]);
			for (std::size_t a = 0; a < in.element.connectivity.size(); ++a) {
				Print(b.value[a]);
				for (int d = 0; d < 3; ++d) Print(b.gradient[a][d]);
			}
		} else if (cmd == "flow" && argc == 4) {
			const auto in = ReadElement(argv[2]);
			Print(iga::IntegrateBoundaryFlow(in.element, std::stoul(argv[3]), in.state));
		} else if (cmd == "traction" && argc == 5) {
			const auto in = ReadElement(argv[2]);
			for (double v : iga::IntegrateBoundaryPressureTraction(in.element, std::stoul(argv[3]), std::stod(argv[4])))
				Print(v);
		} else if (cmd == "surface" && argc == 4) {
			const auto in = ReadElement(argv[2]);
			const int face = std::stoi(argv[3]);
			constexpr int fixed_axis[6] = {2, 1, 0, 1, 0, 2};
			constexpr int varying[6][2] = {{0, 1}, {0, 2}, {1, 2}, {0, 2}, {1, 2}, {0, 1}};
			constexpr double fixed_value[6] = {0.0, 0.0, 1.0, 1.0, 0.0, 1.0};
			for (int qi = 0; qi < 4; ++qi)
				for (int qj = 0; qj < 4; ++qj) {
					std::array<double, 3> c{};
					c[fixed_axis[face]] = fixed_value[face];
					c[varying[face][0]] = iga::kGaussFourPoints[qi];
					c[varying[face][1]] = iga::kGaussFourPoints[qj];
					const auto g = iga::EvaluateElementGeometry(in.element, c);
					Print(iga::SurfaceJacobianMeasure(g, face));
				}
		} else if (cmd == "ntiga" && argc == 3) {
			iga::Database db(argv[2]);
			const auto& h = db.header();
			std::printf("version %u ranks %u elements %llu nodes %llu\n", h.version, h.ranks,
				static_cast<unsigned long long>(h.elements), static_cast<unsigned long long>(h.nodes));
			std::printf("transform %.17g %.17g %.17g %.17g %.17g\n", h.geometry_transform.source_origin[0],
				h.geometry_transform.source_origin[1], h.geometry_transform.source_origin[2],
				h.geometry_transform.source_units_per_normalized_unit, h.geometry_transform.source_length_scale_to_m);
			for (std::uint64_t e = 0; e < h.elements; ++e) {
				const auto el = db.Load(e);
				std::printf("element %llu type %d owner %d nen %zu labels", static_cast<unsigned long long>(el.id),
					el.type, el.owner, el.connectivity.size());
				for (int l : el.boundary_labels) std::printf(" %d", l);
				std::printf("\nconn");
				for (int c : el.connectivity) std::printf(" %d", c);
				std::printf("\nextraction");
				for (const auto& row : el.extraction)
					for (std::size_t p = 0; p < 64; ++p)
						if (row[p] != 0.0) std::printf(" %zu:%.17g", p, row[p]);
				std::printf("\nbezier");
				for (const auto& p : el.bezier_points) std::printf(" %.17g %.17g %.17g", p[0], p[1], p[2]);
				std::printf("\n");
			}
			for (std::uint32_t r = 0; r < h.ranks; ++r) {
				std::printf("rank %u required", r);
				for (auto i : db.RequiredElementIndices(static_cast<std::int32_t>(r)))
					std::printf(" %llu", static_cast<unsigned long long>(i));
				std::printf("\n");
			}
		} else if (cmd == "walltrace" && argc == 4) {
			iga::Database db(argv[2]);
			const auto mesh = iga::ReadLabeledHexMesh(argv[3], db.header().nodes, db.header().elements);
			for (auto n : iga::WallTraceBasis(db, mesh)) std::printf("%d\n", n);
		} else if (cmd == "outlet-config" && argc == 3) {
			const auto config = iga::ReadSimulationConfiguration(argv[2]);
			const auto models = iga::InitializeOutletModels(config, config.equation_systems.at(0));
			auto kind_index = [](iga::FieldBoundaryKind k) {
				return k == iga::FieldBoundaryKind::Resistance ? 0 : k == iga::FieldBoundaryKind::WindkesselRC ? 1 : 2;
			};
			std::printf("%zu\n", models.size());
			for (const auto& m : models)
				std::printf("%d %d %.17g %.17g\n", m.label, kind_index(m.kind), m.pressure, m.capacitor_pressure);
			const auto applied = iga::MaterializeOutletPressures(config, models);
			for (const auto& b : applied.boundaries)
				for (const auto& c : b.conditions)
					if (c.kind == iga::FieldBoundaryKind::PressureTraction)
						std::printf("%d %.17g\n", b.label, c.value.at(0));
		} else if (cmd == "transport-case" && argc == 4) {
			const std::string dir = argv[2];
			const auto nodes = static_cast<std::uint64_t>(std::stoull(argv[3]));
			const auto p = iga::ReadTransportParameters(dir + "/simulation_parameter.txt");
			const auto labels = iga::ReadPointLabels(dir + "/controlmesh.vtk", nodes);
			const auto velocity = iga::ReadVelocity(dir + "/initial_velocityfield.txt", nodes);
			const auto config = iga::ReadCaseConfiguration(dir + "/case_config.json");
			const auto bc = iga::ResolveBoundaryConditions(config, labels, velocity, p);
			for (double v : {p.diffusion, p.vplus, p.vminus, p.kplus, p.kminus, p.detach_plus, p.detach_minus, p.dt,
				static_cast<double>(p.steps), p.n0_bc, p.nplus_bc, p.nminus_bc, p.artificial_diffusion})
				Print(v);
			long long index_sum = 0;
			double n0 = 0.0, nplus = 0.0;
			for (std::size_t i = 0; i < bc.transport_constrained.size(); ++i) {
				if (bc.transport_constrained[i]) index_sum += static_cast<long long>(i);
				n0 += bc.n0[i];
				nplus += bc.nplus[i];
			}
			std::printf("%zu %lld\n", bc.transport_nodes, index_sum);
			Print(n0);
			Print(nplus);
			std::printf("%zu %zu\n", bc.velocity_nodes, bc.pressure_nodes);
		} else if (cmd == "solve-case" && (argc == 4 || argc == 5)) {
			const std::string dir = argv[2];
			iga::Database database(argv[3]);
			const auto config = iga::ReadSimulationConfiguration(dir + "/simulation_config.json");
			const auto name = argc == 5 ? std::string(argv[4]) : iga::FirstLinearTransportSystem(config);
			const auto system = iga::CompileLinearSystem(config, name);
			for (const auto& t : system.terms)
				std::printf("term kind=%d equation=%zu trial=%zu coefficient=%.17g\n", static_cast<int>(t.kind),
					t.equation, t.trial, t.coefficient);
			std::vector<int> supg(system.fields.size(), 0);
			for (const auto& st : system.stabilization)
				supg[system.field_index.at(st.equation)] = st.diffusion_weighting == "supg_test" ? 2
					: st.diffusion_weighting == "consistent" ? 3 : 1;
			const auto tau = iga::SupgTauDiffusivity(system);
			const auto initial = iga::InitialScalarValues(config, system);
			for (std::size_t f = 0; f < system.fields.size(); ++f)
				std::printf("field index=%zu supg=%d tau_diffusivity=%.17g initial=%.17g\n", f, supg[f], tau[f],
					initial[f]);
			const auto labels = iga::ReadPointLabels(dir + "/controlmesh.vtk", database.header().nodes);
			const auto faces = iga::NodeFaceLabels(database);
			const auto bc = iga::ResolveScalarBoundaries(config, system, labels, &faces);
			long long index_sum = 0;
			double value_sum = 0.0;
			for (std::size_t i = 0; i < bc.constrained.size(); ++i)
				if (bc.constrained[i]) {
					index_sum += static_cast<long long>(i);
					value_sum += bc.value[i];
				}
			std::printf("boundary constrained_index_sum=%lld value_sum=%.17g\n", index_sum, value_sum);
			// derived physiology arrays of the state in CASE_DIR/physiology_state.txt (if present)
			std::ifstream state_file(dir + "/physiology_state.txt");
			if (state_file) {
				std::vector<double> values;
				double v = 0.0;
				while (state_file >> v) values.push_back(v);
				for (const auto& a : iga::ComputePhysiologyPointArrays(config.physiology, system.fields, values)) {
					std::printf("derived name=%s", a.name.c_str());
					for (double x : a.values) std::printf(" %.17g", x);
					std::printf("\n");
				}
			}
		} else if (cmd == "outlet" && argc == 12) {
			iga::OutletModelState m;
			m.label = 2;
			m.kind = iga::simulation_detail::ParseFieldBoundaryKind(argv[2]);
			m.resistance = std::stod(argv[3]);
			m.proximal_resistance = std::stod(argv[4]);
			m.distal_resistance = std::stod(argv[5]);
			m.capacitance = std::stod(argv[6]);
			m.reference_pressure = std::stod(argv[7]);
			m.capacitor_pressure = std::stod(argv[8]);
			m.pressure = std::stod(argv[9]);
			const double dt = std::stod(argv[10]), flow = std::stod(argv[11]);
			const auto e = iga::EvaluateOutletModel(m, flow, dt);
			Print(e.pressure);
			Print(e.capacitor_pressure);
			const auto c = iga::EvaluateOutletCoupling({m}, {m.capacitor_pressure}, {flow}, dt);
			Print(c.maximum_pressure_change);
			Print(iga::OutletCouplingTolerance(c));
			auto relaxed = std::vector<iga::OutletModelState>{m};
			iga::RelaxOutletCoupling(relaxed, c);
			Print(relaxed[0].pressure);
			auto committed = std::vector<iga::OutletModelState>{m};
			iga::CommitOutletCoupling(committed, c);
			Print(committed[0].pressure);
			Print(committed[0].capacitor_pressure);
		} else {
			throw std::runtime_error("unknown command or wrong argument count");
		}
		return 0;
	} catch (const std::exception& e) {
		std::fprintf(stderr, "ref_probe: %s\n", e.what());
		return 1;
	}
}
