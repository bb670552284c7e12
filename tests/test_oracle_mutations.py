"""Plant plausible mistakes in the oracle (compile-time switches) and show
that the pinning checks catch each one.  CPU only."""
import pytest

import oracle
import pinning

# planted mistake -> the check that must catch it
MUTATIONS = {
    "ORACLE_MUTATE_GT": "check_scan_equality_threshold",
    "ORACLE_MUTATE_NO_BG_SWAP": "check_pwm_revcompl",
    "ORACLE_MUTATE_LOWER_COUNTS": "check_scan_lowercase_zero",
    "ORACLE_MUTATE_NO_RUN_CHECK": "check_scan_bruteforce",
    "ORACLE_MUTATE_CUTOFF_DOUBLE": "check_cutoff_float_coordinate",
    "ORACLE_MUTATE_HIST_ADD": "check_hist_store_not_add",
    "ORACLE_MUTATE_RUNS_IGNORE_RECORD": "check_fasta_cases",
    "ORACLE_MUTATE_HIST_ROUND": "check_ehist_dyadic_bruteforce",
    "ORACLE_MUTATE_HIST_NO_RUN_CHECK": "check_ehist_total_windows",
    "ORACLE_MUTATE_HIST_FP64": "check_ehist_fp32_order",
    "ORACLE_MUTATE_HIST_RINT": "check_hist_dyadic_exact",
    "ORACLE_MUTATE_HIST_COUNT_DOUBLE": "check_hist_dyadic_exact",
    "ORACLE_MUTATE_JASPAR_ONE_SKIP": "check_load_jaspar_example",
    "ORACLE_MUTATE_PERM_NEEDS_DIGIT": "check_load_cluster_buster",
    "ORACLE_MUTATE_DICT_UPPER_ONLY": "check_dict_closed_form",
    "ORACLE_MUTATE_DICT_FILE_ORDER": "check_dict_closed_form",
}


@pytest.mark.parametrize("define", sorted(MUTATIONS))
def test_mutation_is_caught(define, tmp_path):
    path = oracle.build_variant(define, str(tmp_path))
    mutant = oracle.Oracle(path)
    with pytest.raises(AssertionError):
        getattr(pinning, MUTATIONS[define])(mutant)
