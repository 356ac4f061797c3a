// Output writers (see Output.h).  Fields are pulled from the device first.
#include "Output.h"

#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <string>

using namespace std;

namespace {

template <typename F>
void data_array(ofstream &out, const string &name, int comps, const char *type, F &f) {
    out << "<DataArray Name=\"" << name << "\" NumberOfComponents=\"" << comps << "\" format=\"ascii\" type=\""
        << type << "\">\n";
    out << f;
    out << "</DataArray>\n";
}

ofstream diag_csv;

}  // namespace

// results/fields_%05d.vti, point data: object_id, NodeVol, phi, rho, nd.<sp>, nd-ave.<sp>, ef
void Output::fields(World &world, vector<Species> &species) {
    stringstream name;
    name << "results/fields_" << setfill('0') << setw(5) << world.getTs() << ".vti";
    ofstream out(name.str());
    if (!out.is_open()) {
        cerr << "Could not open " << name.str() << endl;
        return;
    }
    world.updateHost();
    for (Species &sp : species) sp.syncHost();
    double3 x0 = world.getX0(), dh = world.getDh();
    out << "<VTKFile type=\"ImageData\">\n";
    out << "<ImageData Origin=\"" << x0[0] << " " << x0[1] << " " << x0[2] << "\" ";
    out << "Spacing=\"" << dh[0] << " " << dh[1] << " " << dh[2] << "\" ";
    out << "WholeExtent=\"0 " << world.ni - 1 << " 0 " << world.nj - 1 << " 0 " << world.nk - 1 << "\">\n";
    out << "<PointData>\n";
    data_array(out, "object_id", 1, "Int32", world.object_id);
    data_array(out, "NodeVol", 1, "Float64", world.node_vol);
    data_array(out, "phi", 1, "Float64", world.phi);
    data_array(out, "rho", 1, "Float64", world.rho);
    for (Species &sp : species) data_array(out, "nd." + sp.name, 1, "Float64", sp.den);
    for (Species &sp : species) data_array(out, "nd-ave." + sp.name, 1, "Float64", sp.den_ave);
    data_array(out, "ef", 3, "Float64", world.ef);
    out << "</PointData>\n</ImageData>\n</VTKFile>\n";
}

void Output::screenOutput(World &world, vector<Species> &species) {
    cout << "ts: " << world.getTs();
    for (Species &sp : species) cout << setprecision(3) << "\t " << sp.name << ":" << sp.getNp();
    cout << endl;
}

// runtime_diags.csv (ch3/ver2/Output.cpp:97-131): one row per call, in the stream's default
// 6-digit format: ts, time, wall time, then per species the six columns below, then PE and
// the total energy; the header is written by the first call, the rows are flushed every 25
// steps.
void Output::diagOutput(World &world, vector<Species> &species) {
    static const char *const per_species[] = {"mp_count", "real_count", "px", "py", "pz", "KE"};
    if (!diag_csv.is_open()) {
        diag_csv.open("runtime_diags.csv");
        diag_csv << "ts,time,wall_time";
        for (Species &sp : species)
            for (const char *col : per_species) diag_csv << "," << col << "." << sp.name;
        diag_csv << ",PE,E_total" << endl;
    }
    diag_csv << world.getTs() << "," << world.getTime() << "," << world.getWallTime();
    double kinetic = 0;  // summed over the species in their order, then PE added
    for (Species &sp : species) {
        const double ke = sp.getKE();
        double3 p = sp.getMomentum();
        diag_csv << "," << sp.getNp();  // an integer column
        for (double v : {sp.getRealCount(), p[0], p[1], p[2], ke}) diag_csv << "," << v;
        kinetic += ke;
    }
    const double pe = world.getPE();
    diag_csv << "," << pe << "," << (kinetic + pe) << "\n";
    if (world.getTs() % 25 == 0) diag_csv.flush();
}
