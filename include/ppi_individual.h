/*
 * ppi_individual.h — the two types that appear in acc_interpret's signature.
 *
 * They must be layout-identical to PPI's src/individual (individual:11,13) for
 * the drop-in to link against ppi.cc; when that header was included first its
 * guard is honoured and nothing is redefined here.
 */
#ifndef __individual_h
#define __individual_h
#ifndef PPI_INDIVIDUAL_H
#define PPI_INDIVIDUAL_H

typedef bool GENOME_TYPE;                                 /* one allele */
struct Population { GENOME_TYPE** genome; float* fitness; }; /* a population's genomes and fitnesses */

#endif
#endif
