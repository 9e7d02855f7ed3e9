/*
 * serial_chimes_interface -- system-level ChIMES evaluation (C++ API),
 * MI355X implementation.
 *
 * Source-level drop-in for the reference class
 * (serial_interface/src/serial_chimes_interface.h:93-139): construct with the
 * replication flag, init_chimesFF() from a parameter file, calculate() for
 * energy, forces and the 9-component stress (virial / volume).  The whole
 * evaluation -- replication, wrapping, ghost images, cluster enumeration and
 * the 2/3/4-body sums -- runs on the GPU.
 */
#ifndef _serial_chimes_interface_h
#define _serial_chimes_interface_h

#include <string>
#include <vector>

#include "chimesFF.h"

class serial_chimes_interface : public chimesFF
{
public:
    serial_chimes_interface(bool small = true);
    ~serial_chimes_interface();

    bool allow_replication;  // replicate cells shorter than the 2-body cutoff

    void init_chimesFF(string chimesFF_paramfile, int rank);
    void calculate(vector<double> &x_in, vector<double> &y_in, vector<double> &z_in, vector<double> &cella_in,
                   vector<double> &cellb_in, vector<double> &cellc_in, vector<string> &atmtyps, double &energy,
                   vector<vector<double> > &force, vector<double> &stress);
};

#endif
