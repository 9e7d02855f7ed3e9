/*
 * chimesFF -- ChIMES cluster evaluator (C++ API), MI355X implementation.
 *
 * Source-level drop-in for the reference class of the same name
 * (chimesFF/src/chimesFF.h:125-343): the same public members and method
 * signatures.  The evaluation itself runs as HIP kernels on the GPU; the
 * chimes2BTmp / chimes3BTmp / chimes4BTmp scratch classes
 * (chimesFF/src/chimesFF.h:37-117) are kept so existing callers compile, and
 * are not used for the arithmetic.  As in the reference, this header brings
 * namespace std into scope and defines CHDIM.
 */
#ifndef _chimesFF_h
#define _chimesFF_h

#include <string>
#include <vector>

using namespace std;

#define CHDIM 3

class chimes2BTmp
{
public:
    chimes2BTmp(int poly_order) : Tn(poly_order + 1), Tnd(poly_order + 1) {}
    void resize(int poly_order)
    {
        if ((int)Tn.size() < poly_order + 1) Tn.resize(poly_order + 1);
        if ((int)Tnd.size() < poly_order + 1) Tnd.resize(poly_order + 1);
    }
    vector<double> Tn;
    vector<double> Tnd;
};

class chimes3BTmp
{
public:
    chimes3BTmp(int poly_order) { resize(poly_order); }
    void resize(int poly_order)
    {
        for (vector<double> *v : {&Tn_ij, &Tn_ik, &Tn_jk, &Tnd_ij, &Tnd_ik, &Tnd_jk})
            if ((int)v->size() < poly_order + 1) v->resize(poly_order + 1);
    }
    vector<double> Tn_ij, Tn_ik, Tn_jk;
    vector<double> Tnd_ij, Tnd_ik, Tnd_jk;
};

class chimes4BTmp
{
public:
    chimes4BTmp(int poly_order) { resize(poly_order); }
    void resize(int poly_order)
    {
        for (vector<double> *v : {&Tn_ij, &Tn_ik, &Tn_il, &Tn_jk, &Tn_jl, &Tn_kl, &Tnd_ij, &Tnd_ik, &Tnd_il,
                                  &Tnd_jk, &Tnd_jl, &Tnd_kl})
            if ((int)v->size() < poly_order + 1) v->resize(poly_order + 1);
    }
    vector<double> Tn_ij, Tn_ik, Tn_il, Tn_jk, Tn_jl, Tn_kl;
    vector<double> Tnd_ij, Tnd_ik, Tnd_il, Tnd_jk, Tnd_jl, Tnd_kl;
};

struct chimesFF_impl;

class chimesFF
{
public:
    int rank;                 // rank == 0 prints parameter summaries
    int natmtyps;             // number of atom types in the parameter file
    vector<int> poly_orders;  // 2-, 3-, 4-body orders
    vector<string> atmtyps;   // atom type names
    vector<double> masses;    // atom type masses

    chimesFF();
    virtual ~chimesFF();

    void init(int mpi_rank);
    void read_parameters(string paramfile);

    void compute_1B(const int typ_idx, double &energy);
    void compute_2B(const double dx, const vector<double> &dr, const vector<int> typ_idxs, vector<double> &force,
                    vector<double> &stress, double &energy, chimes2BTmp &tmp);
    void compute_2B(const double dx, const vector<double> &dr, const vector<int> typ_idxs, vector<double> &force,
                    vector<double> &stress, double &energy, chimes2BTmp &tmp, double &force_scalar_in);
    void compute_3B(const vector<double> &dx, const vector<double> &dr, const vector<int> &typ_idxs,
                    vector<double> &force, vector<double> &stress, double &energy, chimes3BTmp &tmp);
    void compute_3B(const vector<double> &dx, const vector<double> &dr, const vector<int> &typ_idxs,
                    vector<double> &force, vector<double> &stress, double &energy, chimes3BTmp &tmp,
                    vector<double> &force_scalar_in);
    void compute_4B(const vector<double> &dx, const vector<double> &dr, const vector<int> &typ_idxs,
                    vector<double> &force, vector<double> &stress, double &energy, chimes4BTmp &tmp);
    void compute_4B(const vector<double> &dx, const vector<double> &dr, const vector<int> &typ_idxs,
                    vector<double> &force, vector<double> &stress, double &energy, chimes4BTmp &tmp,
                    vector<double> &force_scalar_in);

    void get_cutoff_2B(vector<vector<double> > &cutoff_2b);
    double max_cutoff_2B(bool silent = false);
    double max_cutoff_3B(bool silent = false);
    double max_cutoff_4B(bool silent = false);

    void set_atomtypes(vector<string> &type_list);
    int get_atom_pair_index(int pair_id);
    void build_pair_int_trip_map();
    void build_pair_int_quad_map();

    int get_badness();
    void reset_badness();

    // Evaluates n explicit clusters of nb = 2, 3 or 4 bodies in one GPU launch
    // (the batched form of compute_2B/3B/4B): dx[n*P] pair distances,
    // dr[n*P*3] pair vectors (x_b - x_a, runtime pair order), typ_idxs[n*nb].
    // energy[n], force[n*nb*3] and stress[n*6] are overwritten.
    void compute_clusters(int nb, int n, const double *dx, const double *dr, const int *typ_idxs, double *energy,
                          double *force, double *stress);

protected:
    chimesFF_impl *impl_;
    int badness;
    void require_model(const char *who);
    void cluster_call(int nb, const double *dx, const double *dr, const int *typ, double *force, double *stress,
                      double &energy, double *last);
};

#endif
